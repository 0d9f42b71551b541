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
//   k_apply                  rho -= drho, u = rou/rho, dt partial minima
// Ghost nodes do not exist at run time: periodic links point at the parent
// (mirror copies are exact copies, mirror_boundaries.F90:263-352).
// Reductions are a fixed-order two-level tree, so results are bitwise
// reproducible run to run.
#include "sv_internal.h"

namespace {

constexpr int BS = 256;   // threads per block of the node kernels

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

struct StageArgs {
  int npfb;
  const int64_t *off;
  const int *kcnt;
  const int *col;
  const double *wx, *wy, *wl;
  const double *gsx, *gsy, *lsum;
  sv_gather G;        // gathered (stage input)
  sv_fields reg_in;   // register 1 input (stage 1: the state)
  sv_fields reg_out;  // register 1 output
  sv_fields e;        // error accumulators
  sv_fields Uout;     // stage 4 result
  sv_gather Gout;     // stages 1-3: next gather buffer
  const double *fx;
  sv_ctrl *ctrl;
  sv_physk pk;
  double *partial;    // stage 4: per-block max error
  int rk1;            // STAGE 0 (right-hand side only): reference stage counter is 1
};

// STAGE 1..4: RK stages; STAGE 0: write the right-hand side into reg_out.
template <int FORM, int STAGE, bool MDIFF>
__global__ __launch_bounds__(BS) void k_stage(StageArgs a) {
  __shared__ double red[BS / SV_WAVE];
  const int bid = sv_xcd_block();
  const int i = bid * BS + threadIdx.x;
  const bool act = i < a.npfb;
  const int ii = act ? i : 0;
  const int64_t b = a.off[ii >> 6] + (ii & 63);
  const int K = act ? a.kcnt[ii] : 0;
  // accumulators
  double rx = 0, ry = 0, ux = 0, uy = 0, vx = 0, vy = 0, lu = 0, lv = 0;
  double axx = 0, ayx = 0, axy = 0, ayy = 0, bxx = 0, byx = 0, bxy = 0, byy = 0;
  double cxx_l = 0, cxy_l = 0, cyy_l = 0;
  double q1x = 0, q1y = 0, q2x = 0, q2y = 0, q3x = 0, q3y = 0;
  for (int k = 0; k < K; k++) {
    const int64_t e = b + (int64_t)k * SV_WAVE;
    const int j = a.col[e];
    const double wxk = a.wx[e], wyk = a.wy[e], wlk = a.wl[e];
    const double roj = a.G.ro[j], uj = a.G.u[j], vj = a.G.v[j];
    rx += roj * wxk; ry += roj * wyk;
    ux += uj * wxk; uy += uj * wyk; lu += uj * wlk;
    vx += vj * wxk; vy += vj * wyk; lv += vj * wlk;
    const double cxxj = a.G.cxx[j], cxyj = a.G.cxy[j], cyyj = a.G.cyy[j];
    if (FORM == SV_FORM_LC) {
      axx += cxxj * wxk;                      // calc_gradient_only(1,cxx)
      axy += cxyj * wxk; ayy += cxyj * wyk;   // calc_gradient(cxy)
      byy += cyyj * wyk;                      // calc_gradient_only(2,cyy)
      const double p1 = a.G.pxx[j], p2 = a.G.pxy[j], p3 = a.G.pyy[j];
      q1x += p1 * wxk; q1y += p1 * wyk;
      q2x += p2 * wxk; q2y += p2 * wyk;
      q3x += p3 * wxk; q3y += p3 * wyk;
    } else {
      axx += cxxj * wxk; ayx += cxxj * wyk;
      axy += cxyj * wxk; ayy += cxyj * wyk;
      bxy += cyyj * wxk; byy += cyyj * wyk;
      if (MDIFF) { cxx_l += cxxj * wlk; cxy_l += cxyj * wlk; cyy_l += cyyj * wlk; }
    }
  }
  (void)bxx; (void)byx;
  double emax_loc = 0.0;
  if (act) {
    const double sx = a.gsx[i], sy = a.gsy[i], sl = a.lsum[i];
    const double ro = a.G.ro[i], u = a.G.u[i], v = a.G.v[i];
    const double cxx = a.G.cxx[i], cxy = a.G.cxy[i], cyy = a.G.cyy[i];
    rx = rx - ro * sx; ry = ry - ro * sy;
    ux = ux - u * sx; uy = uy - u * sy; lu = lu - u * sl;
    vx = vx - v * sx; vy = vy - v * sy; lv = lv - v * sl;
    const sv_physk &P = a.pk;
    // calc_rhs_ro (rhs_lc.F90:127-141)
    const double r_ro = -ro * (ux + vy + 0.0) - (u * rx + v * ry + 0.0 * 0.0);
    // calc_rhs_rovel (rhs_lc.F90:217-272)
    const double tsu = ro * (u * ux + v * uy + 0.0 * 0.0) - u * r_ro;
    const double tsv = ro * (u * vx + v * vy + 0.0 * 0.0) - v * r_ro;
    double dcxx_x, dcxy_x, dcxy_y, dcyy_y;
    dcxx_x = axx - cxx * sx;
    dcxy_x = axy - cxy * sx;
    dcxy_y = ayy - cxy * sy;
    if (FORM == SV_FORM_LC) dcyy_y = byy - cyy * sy;
    else dcyy_y = byy - cyy * sy;
    double fvu = P.coef_solvent * lu;
    double fvv = P.coef_solvent * lv;
    fvu = fvu + P.coef_polymeric * (dcxx_x + dcxy_y + 0.0);
    fvv = fvv + P.coef_polymeric * (dcxy_x + dcyy_y + 0.0);
    double bfu = ro * P.bfx;
    const double bfv = ro * P.bfy;
    if (a.fx) bfu = bfu + a.fx[i];
    const double r_rou = -tsu - P.csq * rx + bfu + fvu;
    const double r_rov = -tsv - P.csq * ry + bfv + fvv;
    double r_xx, r_xy, r_yy;
    if (FORM == SV_FORM_LC) {
      const double pxx = a.G.pxx[i], pxy = a.G.pxy[i], pyy = a.G.pyy[i];
      const bool first = (STAGE == 1 || (STAGE == 0 && a.rk1)) && (a.ctrl->itime == 1);
      rhs_logconf(u, v, ux, uy, vx, vy, pxx, pxy, pyy, q1x - pxx * sx, q1y - pxx * sy,
                  q2x - pxy * sx, q2y - pxy * sy, q3x - pyy * sx, q3y - pyy * sy, first, P.eps,
                  P.lambda, r_xx, r_xy, r_yy);
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
    }
    // ---- RK update (step.F90:291-343, :375-409) ----
    const double rhs[6] = {r_ro, r_rou, r_rov, r_xx, r_xy, r_yy};
    double *const rin[6] = {a.reg_in.ro, a.reg_in.rou, a.reg_in.rov, a.reg_in.pxx,
                            a.reg_in.pxy, a.reg_in.pyy};
    double *const eacc[6] = {a.e.ro, a.e.rou, a.e.rov, a.e.pxx, a.e.pxy, a.e.pyy};
    const sv_ctrl *C = a.ctrl;
    double Un[6];
    if (STAGE == 0) {
      a.reg_out.ro[i] = r_ro; a.reg_out.rou[i] = r_rou; a.reg_out.rov[i] = r_rov;
      a.reg_out.pxx[i] = r_xx; a.reg_out.pxy[i] = r_xy; a.reg_out.pyy[i] = r_yy;
    } else if (STAGE <= 3) {
      const double ra = C->RKa[STAGE - 1], rb = C->RKb[STAGE - 1], rm = C->RKbmbh[STAGE - 1];
      double *const rout[6] = {a.reg_out.ro, a.reg_out.rou, a.reg_out.rov, a.reg_out.pxx,
                               a.reg_out.pxy, a.reg_out.pyy};
#pragma unroll
      for (int f = 0; f < 6; f++) {
        const double r1 = rin[f][i];
        Un[f] = r1 + ra * rhs[f];
        rout[f][i] = r1 + rb * rhs[f];
        eacc[f][i] = (STAGE == 1 ? 0.0 : eacc[f][i]) + rm * rhs[f];
      }
      // derived fields for the next stage (step.F90:350-356)
      const double ooro = 1.0 / Un[0];
      a.Gout.ro[i] = Un[0];
      a.Gout.u[i] = Un[1] * ooro;
      a.Gout.v[i] = Un[2] * ooro;
      if (FORM == SV_FORM_LC) {
        a.Gout.pxx[i] = Un[3];
        a.Gout.pxy[i] = Un[4];
        a.Gout.pyy[i] = Un[5];
        double c1, c2, c3;
        c_from_psi(Un[3], Un[4], Un[5], c1, c2, c3);
        a.Gout.cxx[i] = c1;
        a.Gout.cxy[i] = c2;
        a.Gout.cyy[i] = c3;
      } else {
        a.Gout.cxx[i] = Un[3];
        a.Gout.cxy[i] = Un[4];
        a.Gout.cyy[i] = Un[5];
      }
    } else {
      const double rb = C->RKb[3], rm = C->RKbmbh[3];
      double *const uo[6] = {a.Uout.ro, a.Uout.rou, a.Uout.rov, a.Uout.pxx, a.Uout.pxy,
                             a.Uout.pyy};
      const double nrm[6] = {1.0, P.err_norm_rou, P.err_norm_rou, 1.0, 1.0, 1.0};
#pragma unroll
      for (int f = 0; f < 6; f++) {
        uo[f][i] = rin[f][i] + rb * rhs[f];
        const double ef = eacc[f][i] + rm * rhs[f];
        emax_loc = fmax(emax_loc, fabs(ef) * nrm[f]);
      }
    }
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
  int npfb;
  const int64_t *off;
  const int *kcnt;
  const int *col;
  const double *wh, *hsum, *vol;
  sv_fields U;       // gathered
  sv_fields Uf;      // filtered output
  sv_gather Gout;    // psi (LC) and c of the filtered state
  double rho_char;
  double *partial;   // per-block sum of (ro_f - rho_char)*vol
};

template <int FORM>
__global__ __launch_bounds__(BS) void k_filter(FilterArgs a) {
  __shared__ double red[BS];
  const int bid = sv_xcd_block();
  const int i = bid * BS + threadIdx.x;
  const bool act = i < a.npfb;
  const int ii = act ? i : 0;
  const int64_t b = a.off[ii >> 6] + (ii & 63);
  const int K = act ? a.kcnt[ii] : 0;
  double t0 = 0, t1 = 0, t2 = 0, t3 = 0, t4 = 0, t5 = 0;
  for (int k = 0; k < K; k++) {
    const int64_t e = b + (int64_t)k * SV_WAVE;
    const int j = a.col[e];
    const double w = a.wh[e];
    t0 += a.U.ro[j] * w;
    t1 += a.U.rou[j] * w;
    t2 += a.U.rov[j] * w;
    t3 += a.U.pxx[j] * w;
    t4 += a.U.pxy[j] * w;
    t5 += a.U.pyy[j] * w;
  }
  double m = 0.0;
  if (act) {
    const double hs = a.hsum[i];
    const double f0 = a.U.ro[i], f1 = a.U.rou[i], f2 = a.U.rov[i];
    const double f3 = a.U.pxx[i], f4 = a.U.pxy[i], f5 = a.U.pyy[i];
    const double g0 = f0 + (t0 - f0 * hs), g1 = f1 + (t1 - f1 * hs), g2 = f2 + (t2 - f2 * hs);
    const double g3 = f3 + (t3 - f3 * hs), g4 = f4 + (t4 - f4 * hs), g5 = f5 + (t5 - f5 * hs);
    a.Uf.ro[i] = g0; a.Uf.rou[i] = g1; a.Uf.rov[i] = g2;
    a.Uf.pxx[i] = g3; a.Uf.pxy[i] = g4; a.Uf.pyy[i] = g5;
    if (FORM == SV_FORM_LC) {
      a.Gout.pxx[i] = g3; a.Gout.pxy[i] = g4; a.Gout.pyy[i] = g5;
      double c1, c2, c3;
      c_from_psi(g3, g4, g5, c1, c2, c3);
      a.Gout.cxx[i] = c1; a.Gout.cxy[i] = c2; a.Gout.cyy[i] = c3;
    } else {
      a.Gout.cxx[i] = g3; a.Gout.cxy[i] = g4; a.Gout.cyy[i] = g5;
    }
    m = (g0 - a.rho_char) * a.vol[i];
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.partial[bid] = red[0];
}

// mass correction value, emax and the end-of-step counters
__global__ __launch_bounds__(1024) void k_reduce(int nb, const double *pmass, const double *perr,
                                                 double volsum, sv_ctrl *ctrl) {
  __shared__ double s1[1024], s2[1024];
  double a = 0.0, m = 0.0;
  for (int k = threadIdx.x; k < nb; k += 1024) {
    a += pmass[k];
    m = fmax(m, perr[k]);
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
    ctrl->time = ctrl->time0 + ctrl->RKc[3];
    ctrl->itime = ctrl->itime + 1;               // sunset.F90:99
  }
}

struct ApplyArgs {
  int npfb;
  sv_fields Uf, U;
  sv_gather G;
  const double *s;
  const double *dro;     // &ctrl->dro, or nullptr for no correction
  double cs, visc_total;
  double *pcfl, *pvisc;  // per-block minima for set_tstep
};

// rho -= drho (rhs_lc.F90:887-889); u = rou/rho (step.F90:514-529); set_tstep
// partial minima (step.F90:579-602)
__global__ __launch_bounds__(BS) void k_apply_state(ApplyArgs a) {
  __shared__ double r1[BS], r2[BS];
  const int bid = sv_xcd_block();
  const int i = bid * BS + threadIdx.x;
  double mc = 1.0e10, mv = 1.0e10;
  if (i < a.npfb) {
    const double dro = a.dro ? *a.dro : 0.0;
    const double ro = a.Uf.ro[i] - dro;
    const double rou = a.Uf.rou[i], rov = a.Uf.rov[i];
    a.U.ro[i] = ro; a.U.rou[i] = rou; a.U.rov[i] = rov;
    a.U.pxx[i] = a.Uf.pxx[i]; a.U.pxy[i] = a.Uf.pxy[i]; a.U.pyy[i] = a.Uf.pyy[i];
    const double ooro = 1.0 / ro;
    const double u = rou * ooro, v = rov * ooro;
    a.G.ro[i] = ro; a.G.u[i] = u; a.G.v[i] = v;
    const double um = sqrt(u * u + v * v + 0.0);
    const double si = a.s[i];
    mc = si / (um + a.cs);
    mv = si * si * ro / a.visc_total;
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
__global__ __launch_bounds__(1024) void k_dt(int nb, const double *pcfl, const double *pvisc,
                                             double L_char, sv_ctrl *c) {
  __shared__ double s1[1024], s2[1024];
  double a = 1.0e10, m = 1.0e10;
  for (int k = threadIdx.x; k < nb; k += 1024) {
    a = fmin(a, pcfl[k]);
    m = fmin(m, pvisc[k]);
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

// initial derived fields: u, v and c (LC: from psi) of the state
template <int FORM>
__global__ void k_init_gather(int npfb, sv_fields U, sv_gather G) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npfb) return;
  G.ro[i] = U.ro[i];
  if (FORM == SV_FORM_LC) {
    G.pxx[i] = U.pxx[i]; G.pxy[i] = U.pxy[i]; G.pyy[i] = U.pyy[i];
    double c1, c2, c3;
    c_from_psi(U.pxx[i], U.pxy[i], U.pyy[i], c1, c2, c3);
    G.cxx[i] = c1; G.cxy[i] = c2; G.cyy[i] = c3;
  } else {
    G.cxx[i] = U.pxx[i]; G.cxy[i] = U.pxy[i]; G.cyy[i] = U.pyy[i];
  }
}

StageArgs stage_args(sv_ctx *c, int stage, const sv_gather &Gin, const sv_gather &Gout) {
  StageArgs a{};
  a.npfb = c->npfb;
  a.off = c->sell_off;
  a.kcnt = c->cnt_small;
  a.col = c->scol;
  a.wx = c->swx; a.wy = c->swy; a.wl = c->swl;
  a.gsx = c->gsx; a.gsy = c->gsy; a.lsum = c->lsum;
  a.G = Gin;
  a.reg_in = (stage == 1) ? c->U : c->reg;
  a.reg_out = c->reg;
  a.e = c->eacc;
  a.Uout = c->U;
  a.Gout = Gout;
  a.fx = c->fx;
  a.ctrl = c->ctrl;
  a.pk = c->pk;
  a.partial = c->partial;
  return a;
}

template <int FORM, int STAGE>
void launch_stage(sv_ctx *c, const StageArgs &a, int nb) {
  if (FORM == SV_FORM_DI && c->pk.Mdiff != 0.0)
    k_stage<FORM, STAGE, true><<<nb, BS, 0, c->stream>>>(a);
  else
    k_stage<FORM, STAGE, false><<<nb, BS, 0, c->stream>>>(a);
}

// ev (nullable): SV_NKERNELS+1 events recorded between the launches of one step
template <int FORM>
int step_form(sv_ctx *c, hipEvent_t *ev) {
  const int nb = (c->npfb + BS - 1) / BS;
  double *pm = c->partial, *pe = c->partial + nb, *pc = c->partial + 2 * nb,
         *pv = c->partial + 3 * nb;
  int ke = 0;
  auto mark = [&]() {
    if (ev) (void)hipEventRecord(ev[ke++], c->stream);
  };
  mark();
  k_dt<<<1, 1024, 0, c->stream>>>(nb, pc, pv, c->L_char, c->ctrl);
  mark();
  StageArgs a1 = stage_args(c, 1, c->GA, c->GB);
  launch_stage<FORM, 1>(c, a1, nb);
  mark();
  StageArgs a2 = stage_args(c, 2, c->GB, c->GA);
  launch_stage<FORM, 2>(c, a2, nb);
  mark();
  StageArgs a3 = stage_args(c, 3, c->GA, c->GB);
  launch_stage<FORM, 3>(c, a3, nb);
  mark();
  StageArgs a4 = stage_args(c, 4, c->GB, c->GA);
  a4.partial = pe;
  launch_stage<FORM, 4>(c, a4, nb);
  mark();
  FilterArgs f{};
  f.npfb = c->npfb;
  f.off = c->sell_off;
  f.kcnt = c->cnt;
  f.col = c->scol;
  f.wh = c->swh;
  f.hsum = c->hsum;
  f.vol = c->vol;
  f.U = c->U;
  f.Uf = c->Uf;
  f.Gout = c->GA;
  f.rho_char = c->pk.rho_char;
  f.partial = pm;
  k_filter<FORM><<<nb, BS, 0, c->stream>>>(f);
  mark();
  k_reduce<<<1, 1024, 0, c->stream>>>(nb, pm, pe, c->volsum, c->ctrl);
  mark();
  ApplyArgs ap{};
  ap.npfb = c->npfb;
  ap.Uf = c->Uf;
  ap.U = c->U;
  ap.G = c->GA;
  ap.s = c->s;
  ap.dro = &c->ctrl->dro;
  ap.cs = c->pk.cs;
  ap.visc_total = c->pk.visc_total;
  ap.pcfl = pc;
  ap.pvisc = pv;
  k_apply_state<<<nb, BS, 0, c->stream>>>(ap);
  mark();
  SV_HIP(c, hipGetLastError());
  return 0;
}

}  // namespace

int sv_dev_init_state(sv_ctx *c) {
  const int nb = (c->npfb + BS - 1) / BS;
  if (c->cfg.form == SV_FORM_LC)
    k_init_gather<SV_FORM_LC><<<nb, BS, 0, c->stream>>>(c->npfb, c->U, c->GA);
  else
    k_init_gather<SV_FORM_DI><<<nb, BS, 0, c->stream>>>(c->npfb, c->U, c->GA);
  // u, v and the set_tstep partial minima of the initial state (no correction)
  ApplyArgs ap{};
  ap.npfb = c->npfb;
  ap.Uf = c->U;
  ap.U = c->U;
  ap.G = c->GA;
  ap.s = c->s;
  ap.dro = nullptr;
  ap.cs = c->pk.cs;
  ap.visc_total = c->pk.visc_total;
  ap.pcfl = c->partial + 2 * nb;
  ap.pvisc = c->partial + 3 * nb;
  k_apply_state<<<nb, BS, 0, c->stream>>>(ap);
  SV_HIP(c, hipGetLastError());
  SV_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

int sv_dev_step(sv_ctx *c, hipEvent_t *ev) {
  return c->cfg.form == SV_FORM_LC ? step_form<SV_FORM_LC>(c, ev) : step_form<SV_FORM_DI>(c, ev);
}

// calc_all_rhs on the current state (interior nodes): the stage kernel in its
// right-hand-side-only form, written into the six arrays of out6 (npfb each).
int sv_dev_rhs(sv_ctx *c, int iRKstep, double *out6) {
  const int nb = (c->npfb + BS - 1) / BS;
  const size_t n = c->npfb;
  StageArgs a = stage_args(c, 1, c->GA, c->GB);
  a.rk1 = (iRKstep == 1);
  a.reg_out = sv_fields{out6, out6 + n, out6 + 2 * n, out6 + 3 * n, out6 + 4 * n, out6 + 5 * n,
                        nullptr, nullptr, nullptr};
  if (c->cfg.form == SV_FORM_LC) launch_stage<SV_FORM_LC, 0>(c, a, nb);
  else launch_stage<SV_FORM_DI, 0>(c, a, nb);
  SV_HIP(c, hipGetLastError());
  SV_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

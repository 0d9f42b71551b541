/*
 * sv_oracle_tlm.c -- tangent-linear model (forward-mode derivative) of the
 * direct-form (DI) RK3 step of the oracle, with a fixed time step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the device adjoint): loaded by
 * tests/ alone.
 *
 * For a base state U and a direction dU (ro, rou, rov, cxx, cxy, cyy on real
 * nodes) it advances U exactly as or_step_rk3 does (same operations in the
 * same order, so the base trajectory is bit-identical) and alongside it the
 * derivative of every intermediate value along dU:
 *   - stencil operators are linear: d(grad f) = grad(df), likewise the
 *     Laplacian and the filter (derivatives.F90:26-116, :360-414);
 *   - u = rou/ro: du = (drou - u*dro)/ro (step.F90:514-529);
 *   - the pointwise right-hand sides (rhs_di.F90:127-272 via rhs_lc.F90's
 *     identical calc_rhs_ro/calc_rhs_rovel, and calc_rhs_c rhs_di.F90:417-475
 *     with the oracle's full-gradient reading) by the product rule, term by
 *     term;
 *   - RK stages and registers (step.F90:281-409) are linear combinations;
 *   - the mass correction (rhs_di.F90:667-720, mcorr): dro -= sum(dro*vol)/sum(vol);
 *   - parameters (dWi, dbeta, deps of or_tangent): through the derived
 *     constants of load_data.F90:132-148, visc_solvent = beta*visc_total,
 *     lambda = Wi*L/U, cp = (1-beta)*visc_total/lambda, and the sPTT factor
 *     fr = -(1 - 2 eps + eps*(cxx + cyy + 1))/lambda of calc_rhs_c.
 * The loss is L = (1/T) sum_{t=1..T} mean_i (cxx_i + cyy_i - 2) on the state
 * after each step; or_rollout_tlm returns L and dL = <grad L, dU>.
 *
 * Pinned by central finite differences of or_run (fixed dt) in
 * tests/test_oracle_tlm.py; a dropped or mis-signed term shows as an O(1)
 * relative mismatch there.
 */
#include "sv_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

int or_tangent_alloc(or_tangent *t, int np) {
  memset(t, 0, sizeof(*t));
  t->np = np;
  double **f[8] = {&t->ro, &t->rou, &t->rov, &t->u, &t->v, &t->cxx, &t->cxy, &t->cyy};
  for (int k = 0; k < 8; k++) {
    *f[k] = calloc((size_t)np, sizeof(double));
    if (!*f[k]) return -1;
  }
  return 0;
}

void or_tangent_free(or_tangent *t) {
  free(t->ro); free(t->rou); free(t->rov); free(t->u); free(t->v);
  free(t->cxx); free(t->cxy); free(t->cyy);
  memset(t, 0, sizeof(*t));
}

/* the mirror copies of reapply_mirror_bcs (mirror_boundaries.F90:263-352) are
 * linear: the tangent gets the same copies and sign flips */
static void tangent_mirror(const or_domain *d, or_tangent *t) {
  for (int j = d->npfb; j < d->np; j++) {
    const int i = d->irel[j], vr = d->vrel[j];
    const double su = (vr == 2 || vr == 4) ? -1.0 : 1.0;
    const double sv = (vr == 3 || vr == 4) ? -1.0 : 1.0;
    const double sxy = (vr == 2 || vr == 3) ? -1.0 : 1.0;
    t->ro[j] = t->ro[i];
    t->rou[j] = su * t->rou[i];
    t->rov[j] = sv * t->rov[i];
    t->cxx[j] = t->cxx[i];
    t->cxy[j] = sxy * t->cxy[i];
    t->cyy[j] = t->cyy[i];
  }
}

static void velocity_tangent(const or_domain *d, const or_state *st, or_tangent *t) {
  for (int i = 0; i < d->np; i++) {
    const double ooro = 1.0 / st->ro[i];
    t->u[i] = (t->rou[i] - st->u[i] * t->ro[i]) * ooro;
    t->v[i] = (t->rov[i] - st->v[i] * t->ro[i]) * ooro;
  }
}

static void velocity_base(const or_domain *d, or_state *st) {
  for (int i = 0; i < d->np; i++) {            /* step.F90:514-529 */
    const double ooro = 1.0 / st->ro[i];
    st->u[i] = st->rou[i] * ooro;
    st->v[i] = st->rov[i] * ooro;
  }
}

/* tangent of the DI right-hand side at (st, t); dr[6] sized npfb */
void or_calc_all_rhs_di_tlm(const or_domain *d, const or_phys *p, const or_state *st,
                            const or_tangent *t, double *const dr[6]) {
  const int n = d->npfb;
  enum { RX, RY, UX, UY, VX, VY, LU, LV, AXX, AYX, AXY, AYY, BXY, BYY, LXX, LXY, LYY, NB };
  double *b[NB], *g[NB];
  for (int k = 0; k < NB; k++) {
    b[k] = calloc((size_t)n, sizeof(double));
    g[k] = calloc((size_t)n, sizeof(double));
  }
  /* base derivative values (as or_calc_all_rhs) and their tangents */
  const double *bf[6] = {st->ro, st->u, st->v, st->cxx, st->cxy, st->cyy};
  const double *tf[6] = {t->ro, t->u, t->v, t->cxx, t->cxy, t->cyy};
  const int gx[6] = {RX, UX, VX, AXX, AXY, BXY}, gy[6] = {RY, UY, VY, AYX, AYY, BYY};
  for (int f = 0; f < 6; f++) {
    or_calc_gradient(d, bf[f], b[gx[f]], b[gy[f]]);
    or_calc_gradient(d, tf[f], g[gx[f]], g[gy[f]]);
  }
  or_calc_laplacian(d, st->u, b[LU]);
  or_calc_laplacian(d, st->v, b[LV]);
  or_calc_laplacian(d, t->u, g[LU]);
  or_calc_laplacian(d, t->v, g[LV]);
  const double Md = p->Mdiff;
  if (Md != 0.0) {
    or_calc_laplacian(d, t->cxx, g[LXX]);
    or_calc_laplacian(d, t->cxy, g[LXY]);
    or_calc_laplacian(d, t->cyy, g[LYY]);
  }
  const double csq = p->csq, cs_ = p->visc_solvent, cp = p->visc_polymeric / p->lambda;
  const double eps = p->epsPTT, lam = p->lambda;
  /* parameter tangents of the derived constants */
  const double dlam = t->dWi * p->L_char / p->U_char;
  const double dcs = t->dbeta * p->visc_total;
  const double dcp = -t->dbeta * p->visc_total / lam - cp * dlam / lam;
  const double bfx = p->grav[0] + p->drive[0], bfy = p->grav[1] + p->drive[1];
#pragma omp parallel for
  for (int i = 0; i < n; i++) {
    const double ro = st->ro[i], u = st->u[i], v = st->v[i];
    const double cxx = st->cxx[i], cxy = st->cxy[i], cyy = st->cyy[i];
    const double dro = t->ro[i], du = t->u[i], dv = t->v[i];
    const double dcxx = t->cxx[i], dcxy = t->cxy[i], dcyy = t->cyy[i];
    const double rx = b[RX][i], ry = b[RY][i], ux = b[UX][i], uy = b[UY][i];
    const double vx = b[VX][i], vy = b[VY][i];
    const double drx = g[RX][i], dry = g[RY][i], dux = g[UX][i], duy = g[UY][i];
    const double dvx = g[VX][i], dvy = g[VY][i], dlu = g[LU][i], dlv = g[LV][i];
    const double axx = b[AXX][i], ayx = b[AYX][i], axy = b[AXY][i], ayy = b[AYY][i];
    const double bxy = b[BXY][i], byy = b[BYY][i];
    const double daxx = g[AXX][i], dayx = g[AYX][i], daxy = g[AXY][i], dayy = g[AYY][i];
    const double dbxy = g[BXY][i], dbyy = g[BYY][i];
    /* calc_rhs_ro: r = -ro*(ux+vy) - (u*rx + v*ry) */
    const double r_ro = -ro * (ux + vy) - (u * rx + v * ry);
    const double d_ro = -dro * (ux + vy) - ro * (dux + dvy) - (du * rx + u * drx + dv * ry + v * dry);
    /* calc_rhs_rovel */
    const double d_tsu = dro * (u * ux + v * uy) + ro * (du * ux + u * dux + dv * uy + v * duy)
                         - du * r_ro - u * d_ro;
    const double d_tsv = dro * (u * vx + v * vy) + ro * (du * vx + u * dvx + dv * vy + v * dvy)
                         - dv * r_ro - v * d_ro;
    const double d_fvu = cs_ * dlu + cp * (daxx + dayy) + dcs * b[LU][i] + dcp * (axx + ayy);
    const double d_fvv = cs_ * dlv + cp * (daxy + dbyy) + dcs * b[LV][i] + dcp * (axy + byy);
    dr[0][i] = d_ro;
    dr[1][i] = -d_tsu - csq * drx + dro * bfx + d_fvu;
    dr[2][i] = -d_tsv - csq * dry + dro * bfy + d_fvv;
    /* calc_rhs_c (DI) */
    const double d_adxx = -(du * axx + u * daxx + dv * ayx + v * dayx);
    const double d_adxy = -(du * axy + u * daxy + dv * ayy + v * dayy);
    const double d_adyy = -(du * bxy + u * dbxy + dv * byy + v * dbyy);
    const double d_ucxx = 2.0 * (dux * cxx + ux * dcxx) + 2.0 * (duy * cxy + uy * dcxy);
    const double d_ucxy = dux * cxy + ux * dcxy + duy * cyy + uy * dcyy + dvx * cxx + vx * dcxx
                          + dvy * cxy + vy * dcxy;
    const double d_ucyy = 2.0 * (dvx * cxy + vx * dcxy) + 2.0 * (dvy * cyy + vy * dcyy);
    const double fr = -(1.0 - eps * 2.0 + eps * (cxx + cyy + 1.0)) / lam;
    const double dfr = -eps * (dcxx + dcyy) / lam - (cxx + cyy - 1.0) * t->deps / lam
                       - fr * dlam / lam;
    double mxx = 0, mxy = 0, myy = 0;
    if (Md != 0.0) { mxx = Md * g[LXX][i]; mxy = Md * g[LXY][i]; myy = Md * g[LYY][i]; }
    dr[3][i] = d_adxx + d_ucxx + dfr * (cxx - 1.0) + fr * dcxx + mxx;
    dr[4][i] = d_adxy + d_ucxy + dfr * cxy + fr * dcxy + mxy;
    dr[5][i] = d_adyy + d_ucyy + dfr * (cyy - 1.0) + fr * dcyy + myy;
  }
  for (int k = 0; k < NB; k++) { free(b[k]); free(g[k]); }
}

/* mean_i (cxx + cyy - 2) over real nodes */
double or_loss_trace(const or_domain *d, const or_state *st) {
  double s = 0.0;
  for (int i = 0; i < d->npfb; i++) s += st->cxx[i] + st->cyy[i] - 2.0;
  return s / d->npfb;
}

static double loss_trace_tangent(const or_domain *d, const or_tangent *t) {
  double s = 0.0;
  for (int i = 0; i < d->npfb; i++) s += t->cxx[i] + t->cyy[i];
  return s / d->npfb;
}

/* One DI step (fixed dt) of base and tangent; the base follows or_step_rk3. */
int or_step_rk3_di_tlm(const or_domain *d, const or_phys *p, or_state *st, or_tangent *t) {
  if (p->form != OR_FORM_DI) return -1;
  const int n = d->npfb;
  const double a[3] = {11847461282814.0 / 36547543011857.0, 3943225443063.0 / 7078155732230.0,
                       -346793006927.0 / 4029903576067.0};
  const double b[4] = {1017324711453.0 / 9774461848756.0, 8237718856693.0 / 13685301971492.0,
                       57731312506979.0 / 19404895981398.0,
                       -101169746363290.0 / 37734290219643.0};
  double RKa[3], RKb[4];
  for (int k = 0; k < 3; k++) RKa[k] = st->dt * a[k];
  for (int k = 0; k < 4; k++) RKb[k] = st->dt * b[k];
  double *reg[6], *rhs[6], *treg[6], *trhs[6];
  for (int f = 0; f < 6; f++) {
    reg[f] = malloc(sizeof(double) * n); rhs[f] = malloc(sizeof(double) * n);
    treg[f] = malloc(sizeof(double) * n); trhs[f] = malloc(sizeof(double) * n);
  }
  double *U[6] = {st->ro, st->rou, st->rov, st->cxx, st->cxy, st->cyy};
  double *T[6] = {t->ro, t->rou, t->rov, t->cxx, t->cxy, t->cyy};
  for (int f = 0; f < 6; f++) {
    memcpy(reg[f], U[f], sizeof(double) * n);
    memcpy(treg[f], T[f], sizeof(double) * n);
  }
  for (int k = 0; k < 4; k++) {
    /* base: the same calls as or_step_rk3 (the error accumulator does not feed
     * the state at a fixed dt) */
    or_calc_all_rhs(d, p, st, k + 1, rhs[0], rhs[1], rhs[2], rhs[3], rhs[4], rhs[5]);
    or_calc_all_rhs_di_tlm(d, p, st, t, trhs);
    if (k < 3) {
      for (int i = 0; i < n; i++)
        for (int f = 0; f < 6; f++) {
          U[f][i] = reg[f][i] + RKa[k] * rhs[f][i];
          reg[f][i] = reg[f][i] + RKb[k] * rhs[f][i];
          T[f][i] = treg[f][i] + RKa[k] * trhs[f][i];
          treg[f][i] = treg[f][i] + RKb[k] * trhs[f][i];
        }
      or_reapply_mirror_bcs(d, p, st);
      tangent_mirror(d, t);
      velocity_base(d, st);
      velocity_tangent(d, st, t);
    } else {
      for (int i = 0; i < n; i++)
        for (int f = 0; f < 6; f++) {
          U[f][i] = reg[f][i] + RKb[3] * rhs[f][i];
          T[f][i] = treg[f][i] + RKb[3] * trhs[f][i];
        }
    }
  }
  for (int f = 0; f < 6; f++) { free(reg[f]); free(rhs[f]); free(treg[f]); free(trhs[f]); }
  or_reapply_mirror_bcs(d, p, st);
  tangent_mirror(d, t);
  or_filter_variables(d, p, st);                /* filter + mcorr of the base */
  double *tf[6] = {t->ro, t->rou, t->rov, t->cxx, t->cxy, t->cyy};
  for (int f = 0; f < 6; f++) or_calc_filtered_var(d, tf[f]);
  double tm = 0.0, tv = 0.0;
  for (int i = 0; i < n; i++) { tm += t->ro[i] * d->vol[i]; tv += d->vol[i]; }
  const double tdro = tm / tv;
  for (int i = 0; i < n; i++) t->ro[i] -= tdro;
  or_reapply_mirror_bcs(d, p, st);
  tangent_mirror(d, t);
  velocity_base(d, st);
  velocity_tangent(d, st, t);
  st->time += st->dt;
  return 0;
}

/* Set the tangent of the real-node state (ro, rou, rov, cxx, cxy, cyy), fill
 * ghosts and velocity tangents. */
void or_tangent_set(const or_domain *d, const or_state *st, or_tangent *t, const double *dro,
                    const double *drou, const double *drov, const double *dcxx,
                    const double *dcxy, const double *dcyy) {
  for (int i = 0; i < d->npfb; i++) {
    t->ro[i] = dro[i]; t->rou[i] = drou[i]; t->rov[i] = drov[i];
    t->cxx[i] = dcxx[i]; t->cxy[i] = dcxy[i]; t->cyy[i] = dcyy[i];
  }
  tangent_mirror(d, t);
  velocity_tangent(d, st, t);
}

/* nsteps of the fixed-dt DI loop with the tangent; loss and its derivative */
int or_rollout_tlm(const or_domain *d, const or_phys *p, or_state *st, or_tangent *t,
                   int nsteps, double *loss, double *dloss) {
  double L = 0.0, dL = 0.0;
  for (int s = 0; s < nsteps; s++) {
    if (or_step_rk3_di_tlm(d, p, st, t)) return -1;
    st->itime++;
    L += or_loss_trace(d, st);
    dL += loss_trace_tangent(d, t);
  }
  if (loss) *loss = L / nsteps;
  if (dloss) *dloss = dL / nsteps;
  return 0;
}

/*
 * sv_oracle_bound.c -- CPU restatement of the boundary discretisation:
 * boundary node lists and FD lineage (setup_domain.F90:158-202, :314-326) and
 * calc_boundary_weights (labf.F90:505-1039).
 *
 * TEST INFRASTRUCTURE ONLY (the checker), like sv_oracle.c.
 *
 * Boundary node i (type 0 wall, 1 inflow, 2 outflow) is followed in index order
 * by its four FD rows i+1..i+4 (types -1..-4) at distances s, 2s, 3s, 4s along
 * its unit normal, as the IPART reader generates them (setup_domain.F90:179-193).
 * Operators on a boundary node: 5-point one-sided differences along the normal,
 * 1D LABFM (6th order, Hermite ABFs times the Wendland C2 derivative kernel fac)
 * along the tangent.  Rows 1 and 2: centred/one-sided 5-point differences along
 * the normal, 1D LABFM along a tangent rotated by the tangential gradient of s.
 * The 6x6 systems are solved with or_svd_solve, the restated svd_solve.
 */
#include "sv_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double PI_B = 3.141592653589793238462643383279502884197;

int or_set_boundaries(or_domain *d, const double *rnx, const double *rny) {
  free(d->blist); free(d->fdp); free(d->rnx); free(d->rny); free(d->wbt); free(d->wbtsum);
  d->blist = NULL;
  d->wbt = d->wbtsum = NULL;
  d->nb = 0;
  d->fdp = malloc(sizeof(int) * d->cap);
  d->rnx = calloc(d->cap, sizeof(double));
  d->rny = calloc(d->cap, sizeof(double));
  if (!d->fdp || !d->rnx || !d->rny) return -1;
  int nb = 0;
  for (int i = 0; i < d->npfb; i++) {
    d->rnx[i] = rnx ? rnx[i] : 0.0;
    d->rny[i] = rny ? rny[i] : 0.0;
    d->fdp[i] = -1;
    const int t = d->node_type[i];
    if (t >= 0 && t <= 2) nb++;
  }
  for (int k = d->npfb; k < d->np; k++) d->fdp[k] = -1;
  d->blist = malloc(sizeof(int) * (nb ? nb : 1));
  if (!d->blist) return -1;
  for (int i = 0; i < d->npfb; i++) {
    const int t = d->node_type[i];
    if (t >= 0 && t <= 2) {
      if (i + 4 >= d->npfb) return -1;
      for (int j = 1; j <= 4; j++)
        if (d->node_type[i + j] != -j) return -1;
      d->blist[d->nb++] = i;
    } else if (t < 0) {
      if (t < -4 || i + t < 0) return -1;
      const int p = i + t;
      if (d->node_type[p] < 0 || d->node_type[p] > 2) return -1;
      d->fdp[i] = p;
      d->rnx[i] = d->rnx[p];                   /* setup_domain.F90:189 */
      d->rny[i] = d->rny[p];
    }
  }
  for (int k = d->npfb; k < d->np; k++) {      /* mirror_boundaries.F90:253 */
    d->rnx[k] = d->rnx[d->irel[k]];
    d->rny[k] = d->rny[d->irel[k]];
  }
  return 0;
}

/* rbfs.F90:58-68 (kernel 1): the Wendland C2 derivative kernel, h3 = h0^3 */
static double fac_b(double qq, double h3) {
  const double qq2 = qq * qq, qq3 = qq * qq2, qq4 = qq * qq3;
  double f = 0.0;
  if (qq <= 2.0 && qq > 0.0)
    f = (7.0 / (4.0 * PI_B * h3)) * (0.625 * qq4 - 3.75 * qq3 + 7.5 * qq2 - 5.0 * qq);
  return f;
}

/* monomials1D, labf.F90:1041-1048 */
static void mono1(double z, double *c) {
  const double z2 = z * z;
  c[0] = z; c[1] = 0.5 * z2; c[2] = (1.0 / 6.0) * z * z2; c[3] = z2 * z2 / 24.0;
  c[4] = z2 * z2 * z / 120.0; c[5] = z2 * z2 * z2 / 720.0;
}

/* abfs1D with Hermite ABFs, labf.F90:1050-1077, Hermite1..6 (:2329-2368) */
static void abf1(double z, double ff1, double *c) {
  const double z2 = z * z, z3 = z * z2;
  c[0] = ff1 * z;
  c[1] = ff1 * (z * z - 1.0);
  c[2] = ff1 * (1.0 / sqrt(2.0)) * (z * z * z - 3.0 * z);
  c[3] = ff1 * (z * z * z * z - 6.0 * z * z + 3.0);
  c[4] = ff1 * (z2 * z3 - 10.0 * z3 + 15.0 * z);
  c[5] = ff1 * (z3 * z3 - 15.0 * z2 * z2 + 45.0 * z2 - 15.0);
}

/* Tangential coordinate of j relative to i, scaled by hh (labf.F90:704-712) */
static double tcoord(const or_domain *d, int i, int j, double xt, double yt) {
  const double rx = d->x[i] - d->x[j], ry = d->y[i] - d->y[j];
  double x = -(rx * xt + ry * yt);
  if (i == j) x = 0.0;
  return x;
}

/* The 1D system over the same-layer neighbours of i and its solutions for the
 * right-hand sides e_{r1}, e_{r2}, hs*e_{r3} (0-based r; r3 < 0: none). */
static void solve_1d(const or_domain *d, int i, double xt, double yt, double hh, double h3,
                     int r1, int r2, int r3, double hs, double b1[6], double b2[6],
                     double b3[6]) {
  double A[36];
  memset(A, 0, sizeof(A));
  const size_t base = (size_t)i * d->nplink;
  for (int k = 0; k < d->ij_count[i]; k++) {
    const int j = d->ij_link[base + k];
    if (d->node_type[j] != d->node_type[i]) continue;
    double x = tcoord(d, i, j, xt, yt);
    double xv[6], g[6];
    mono1(x / hh, xv);
    const double ff1 = fac_b(fabs(x / hh), h3);
    x = x / hh;
    abf1(x, ff1, g);
    for (int a = 0; a < 6; a++)           /* amatt(i1,i2) += xvec(i1)*gvec(i2) */
      for (int c = 0; c < 6; c++) A[c * 6 + a] += xv[a] * g[c];
  }
  memset(b1, 0, sizeof(double) * 6); b1[r1] = 1.0;
  memset(b2, 0, sizeof(double) * 6); b2[r2] = 1.0;
  or_svd_solve(6, A, b1);
  or_svd_solve(6, A, b2);
  if (r3 >= 0) {
    memset(b3, 0, sizeof(double) * 6); b3[r3] = hs;
    or_svd_solve(6, A, b3);
  }
}

/* dot(b, abfs1D(x/hh)) for neighbour j of i */
static double w1d(const or_domain *d, int i, int j, double xt, double yt, double hh,
                  double h3, const double b[6]) {
  double x = tcoord(d, i, j, xt, yt);
  const double ff1 = fac_b(fabs(x / hh), h3);
  x = x / hh;
  double g[6];
  abf1(x, ff1, g);
  double s = 0.0;
  for (int a = 0; a < 6; a++) s += b[a] * g[a];
  return s;
}

int or_boundary_weights(or_domain *d) {
  const int nb = d->nb, NPL = d->nplink;
  if (nb == 0) return 0;
  const double h3 = d->h2 * d->h0;
  double *wb2 = calloc((size_t)nb * NPL * 2, sizeof(double));   /* ij_wb_grad2(1:2,:,jj) */
  double *ooR = calloc(nb, sizeof(double));
  int *tmp = malloc(sizeof(int) * NPL);
  if (!wb2 || !ooR || !tmp) return -1;
#define WG(i, k, c) d->w_grad[(((size_t)(i) * NPL) + (k)) * 2 + (c)]
#define WL(i, k) d->w_lap[(size_t)(i) * NPL + (k)]
#define WH(i, k) d->w_hyp[(size_t)(i) * NPL + (k)]
#define LNK(i, k) d->ij_link[(size_t)(i) * NPL + (k)]
#define B2(jj, k, c) wb2[(((size_t)(jj) * NPL) + (k)) * 2 + (c)]
  /* Preface: a boundary node keeps its same-type neighbours and its FD rows
   * (labf.F90:541-566) */
  for (int jj = 0; jj < nb; jj++) {
    const int i = d->blist[jj], kk = d->node_type[i];
    const int n = d->ij_count[i];
    memcpy(tmp, &LNK(i, 0), sizeof(int) * n);
    for (int k = 0; k < NPL; k++) LNK(i, k) = 0;
    int c = 0;
    for (int k = 0; k < n; k++) {
      const int j = tmp[k];
      if (d->node_type[j] == kk) LNK(i, c++) = j;
      if (j < d->npfb && d->fdp[j] == i) LNK(i, c++) = j;
    }
    d->ij_count[i] = c;
    d->ij_count_small[i] = c;
  }
  /* Normal operators on the boundary node (:568-612) */
  for (int jj = 0; jj < nb; jj++) {
    const int i = d->blist[jj];
    const double dx = d->s[i], dx2 = dx * dx;
    for (int k = 0; k < NPL; k++) { WG(i, k, 0) = 0; WG(i, k, 1) = 0; WH(i, k) = 0; }
    for (int k = 0; k < d->ij_count[i]; k++) {
      const int j = LNK(i, k);
      if (j >= d->npfb) continue;
      if (d->fdp[j] != i && j != i) continue;
      if (j == i) { WG(i, k, 0) = 0.0; B2(jj, k, 0) = 0.0; continue; }
      switch (d->node_type[j]) {
        case -1: WG(i, k, 0) = 4.0 / dx; B2(jj, k, 0) = -104.0 / 12.0 / dx2; break;
        case -2: WG(i, k, 0) = -3.0 / dx; B2(jj, k, 0) = 114.0 / 12.0 / dx2; break;
        case -3: WG(i, k, 0) = (4.0 / 3.0) / dx; B2(jj, k, 0) = -56.0 / 12.0 / dx2; break;
        case -4: WG(i, k, 0) = -0.25 / dx; B2(jj, k, 0) = 11.0 / 12.0 / dx2; break;
        default: break;
      }
    }
  }
  /* Row 1 normal operators (:615-645) */
  for (int jj = 0; jj < nb; jj++) {
    const int i = d->blist[jj] + 1, par = d->fdp[i];
    const double dx = d->s[i], dx2 = dx * dx;
    for (int k = 0; k < NPL; k++) { WG(i, k, 0) = 0; WG(i, k, 1) = 0; WL(i, k) = 0; }
    for (int k = 0; k < d->ij_count[i]; k++) {
      const int j = LNK(i, k);
      if (j >= d->npfb) continue;
      const int sib = (d->fdp[j] == par);
      if (j == par) { WG(i, k, 0) = -0.25 / dx; WL(i, k) = 11.0 / 12.0 / dx2; }
      if (j == i) { WG(i, k, 0) = 0.0; WL(i, k) = 0.0; }
      if (d->node_type[j] == -2 && sib) { WG(i, k, 0) = 1.5 / dx; WL(i, k) = 6.0 / 12.0 / dx2; }
      if (d->node_type[j] == -3 && sib) { WG(i, k, 0) = -0.5 / dx; WL(i, k) = 4.0 / 12.0 / dx2; }
      if (d->node_type[j] == -4 && sib) { WG(i, k, 0) = 1.0 / 12.0 / dx; WL(i, k) = -1.0 / 12.0 / dx2; }
    }
  }
  /* Row 2 normal operators (:647-677) */
  for (int jj = 0; jj < nb; jj++) {
    const int i = d->blist[jj] + 2, par = d->fdp[i];
    const double dx = d->s[i], dx2 = dx * dx;
    for (int k = 0; k < NPL; k++) { WG(i, k, 0) = 0; WG(i, k, 1) = 0; WL(i, k) = 0; }
    for (int k = 0; k < d->ij_count[i]; k++) {
      const int j = LNK(i, k);
      if (j >= d->npfb) continue;
      const int sib = (d->fdp[j] == par);
      if (j == par) { WG(i, k, 0) = 1.0 / 12.0 / dx; WL(i, k) = -1.0 / 12.0 / dx2; }
      if (d->node_type[j] == -1 && sib) { WG(i, k, 0) = -8.0 / 12.0 / dx; WL(i, k) = 16.0 / 12.0 / dx2; }
      if (j == i) { WG(i, k, 0) = 0.0 / 12.0 / dx; WL(i, k) = 0.0 / 12.0 / dx2; }
      if (d->node_type[j] == -3 && sib) { WG(i, k, 0) = 8.0 / 12.0 / dx; WL(i, k) = 16.0 / 12.0 / dx2; }
      if (d->node_type[j] == -4 && sib) { WG(i, k, 0) = -1.0 / 12.0 / dx; WL(i, k) = -1.0 / 12.0 / dx2; }
    }
  }
  /* Transverse operators on the boundary node (:692-758) */
  for (int jj = 0; jj < nb; jj++) {
    const int i = d->blist[jj];
    const double xt = -d->rny[i], yt = d->rnx[i], hh = d->h[i];
    double bt[6], btt[6], bth[6];
    solve_1d(d, i, xt, yt, hh, h3, 0, 1, 3, -1.0, bt, btt, bth);
    for (int k = 0; k < d->ij_count[i]; k++) {
      const int j = LNK(i, k);
      if (d->node_type[j] != d->node_type[i]) continue;
      WG(i, k, 1) = w1d(d, i, j, xt, yt, hh, h3, bt) / hh;
      B2(jj, k, 1) = w1d(d, i, j, xt, yt, hh, h3, btt) / hh / hh;
      WH(i, k) = WH(i, k) + w1d(d, i, j, xt, yt, hh, h3, bth);
    }
  }
  /* Transverse operators on rows 1 and 2 (:761-923): the tangent is rotated by
   * -atan of the parent's tangential derivative of s */
  for (int row = 1; row <= 2; row++) {
    for (int jj = 0; jj < nb; jj++) {
      const int i = d->blist[jj] + row, ii = d->fdp[i];
      const double hh = d->h[i];
      double grads = 0.0;
      for (int k = 0; k < d->ij_count[ii]; k++) {
        const int j = LNK(ii, k);
        grads = grads + (d->s[j] - d->s[ii]) * WG(ii, k, 1);
      }
      grads = -atan(grads);
      const double xt = cos(grads) * (-d->rny[i]) - sin(grads) * d->rnx[i];
      const double yt = sin(grads) * (-d->rny[i]) + cos(grads) * d->rnx[i];
      double bt[6], btt[6], bth[6];
      solve_1d(d, i, xt, yt, hh, h3, 0, 1, -1, 0.0, bt, btt, bth);
      for (int k = 0; k < d->ij_count[i]; k++) {
        const int j = LNK(i, k);
        if (d->node_type[j] != d->node_type[i]) continue;
        WG(i, k, 1) = w1d(d, i, j, xt, yt, hh, h3, bt) / hh;
        WL(i, k) = WL(i, k) + w1d(d, i, j, xt, yt, hh, h3, btt) / hh / hh;
      }
      for (int k = 0; k < d->ij_count[i]; k++)
        WG(i, k, 1) = (1.0 / cos(grads)) * WG(i, k, 1) - tan(grads) * WG(i, k, 0);
    }
  }
  /* Wall curvature 1/R from the tangential second derivative of the normal
   * (:929-945) */
  for (int jj = 0; jj < nb; jj++) {
    const int i = d->blist[jj];
    if (d->node_type[i] != 0) continue;
    double tt = 0.0;
    for (int k = 0; k < d->ij_count[i]; k++) {
      const int j = LNK(i, k);
      if (d->node_type[j] != d->node_type[i]) continue;
      const double tn = (d->rnx[j] * d->rnx[i] + d->rny[j] * d->rny[i]) - 1.0;
      tt = tt + B2(jj, k, 1) * tn;
    }
    ooR[jj] = sqrt(fabs(tt));
  }
  /* Outflow: normal-tangent frame to x-y by reversing (:949-954) */
  for (int jj = 0; jj < nb; jj++) {
    const int i = d->blist[jj];
    if (d->node_type[i] == 2)
      for (int k = 0; k < NPL; k++) { WG(i, k, 0) = -WG(i, k, 0); WG(i, k, 1) = -WG(i, k, 1); }
  }
  /* Divergence term of the Laplacian on wall rows 1 and 2 (:957-966) */
  for (int jj = 0; jj < nb; jj++) {
    int i = d->blist[jj] + 1;
    if (d->node_type[i - 1] == 0)
      for (int k = 0; k < NPL; k++)
        WL(i, k) = WL(i, k) + WG(i, k, 0) * ooR[jj] / (1.0 + ooR[jj] * d->s[i]);
    i = d->blist[jj] + 2;
    if (d->node_type[i - 2] == 0)
      for (int k = 0; k < NPL; k++)
        WL(i, k) = WL(i, k) + WG(i, k, 0) * ooR[jj] / (1.0 + 2.0 * ooR[jj] * d->s[i]);
  }
  /* Walls: curvature term of d2/dn2 and the Laplacian (:970-980) */
  for (int jj = 0; jj < nb; jj++) {
    const int i = d->blist[jj];
    if (d->node_type[i] != 0) continue;
    for (int k = 0; k < d->ij_count[i]; k++) {
      B2(jj, k, 0) = B2(jj, k, 0) + WG(i, k, 0) * ooR[jj];
      WL(i, k) = B2(jj, k, 0) + B2(jj, k, 1);
    }
  }
  /* keep the tangential second-derivative weights of every boundary node */
  free(d->wbt); free(d->wbtsum);
  d->wbt = calloc((size_t)nb * NPL, sizeof(double));
  d->wbtsum = calloc(nb, sizeof(double));
  if (!d->wbt || !d->wbtsum) { free(wb2); free(ooR); free(tmp); return -1; }
  for (int jj = 0; jj < nb; jj++)
    for (int k = 0; k < NPL; k++) d->wbt[(size_t)jj * NPL + k] = B2(jj, k, 1);
  /* Rows 1 and 2: normal-tangent weights to the x-y frame (:984-1000) */
  for (int jj = 0; jj < nb; jj++) {
    const int ib = d->blist[jj];
    const double xn = d->rnx[ib], yn = d->rny[ib];
    for (int row = 1; row <= 2; row++) {
      const int i = ib + row;
      for (int k = 0; k < d->ij_count[i]; k++) {
        const double gn = WG(i, k, 0), gt = WG(i, k, 1);
        WG(i, k, 0) = xn * gn + (-yn) * gt;
        WG(i, k, 1) = yn * gn + xn * gt;
      }
    }
  }
#undef WG
#undef WL
#undef WH
#undef LNK
#undef B2
  free(wb2); free(ooR); free(tmp);
  return 0;
}

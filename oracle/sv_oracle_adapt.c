/*
 * sv_oracle_adapt.c -- stencil adaptation (adapt_stencils, labf.F90:1079-1416;
 * adapt_stencils_main, :1418-1721) for the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (the checker), like sv_oracle.c.
 *
 * The adaptation reduces each interior node's h by 2% while the residual of
 * the SVD solution of its Laplacian moment system stays below res_tol/hh and
 * the d/dx, d/dy, Laplacian operators do not amplify any of 33 wavenumbers up
 * to (just past) the node's Nyquist wavenumber.  The residual test sits at the
 * level of the SVD's own rounding, so the decisions are reproduced only by an
 * SVD with the reference's operation order: or_svd_gr below is the
 * Golub-Reinsch algorithm (Householder bidiagonalisation, accumulation of the
 * right and left transformations, implicit-shift QR on the bidiagonal form,
 * back substitution skipping zero singular values) in the published
 * Numerical Recipes formulation that svdlib.F90:22-306 follows, with the same
 * sequence of floating-point operations (compiled with -ffp-contract=off), and
 * gr_pythag the overflow-safe hypotenuse of svdlib.F90:308-327.  The moment
 * matrices use or_monomials, or_abfs and or_wab (pow(t, 4.0), as the
 * reference's real power compiles), which evaluate the reference's
 * expressions in its order.
 *
 * Constants as the reference declares them: the reduction factor 0.98 and the
 * limit log(0.6) are single-precision literals assigned to double
 * (labf.F90:1129, :1306); res_tol = c * n^4 * eps / m with c = 1e-3, 5e-3,
 * 3e-2 for orders 4, 6, 8 (:1130-1145); amp_tol = 1.0001; 32 wavenumbers.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sv_oracle.h"

static const double PI_A = 3.141592653589793238462643383279502884197;

static double gr_pythag(double a, double b) {
  const double aa = fabs(a), ab = fabs(b);
  if (aa > ab) {
    const double r = ab / aa;
    return aa * sqrt(1.0 + r * r);
  }
  if (ab == 0.0) return 0.0;
  const double r = aa / ab;
  return ab * sqrt(1.0 + r * r);
}

/* Fortran SIGN(a, b): |a| with the sign of b */
static double f_sign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

#define UM(r, c) U[(size_t)(c) * n + (r)]
#define VM(r, c) V[(size_t)(c) * n + (r)]

/* Golub-Reinsch SVD of the n x n column-major matrix in U (overwritten by U);
 * W singular values, V right vectors (n x n), rv1 work (n).  0-based indices
 * throughout; the loops run in the order of the 1-based original.  Returns 0,
 * or -1 after 30 QR iterations on one singular value without convergence. */
int or_svd_gr(int n, double *U, double *W, double *V, double *rv1) {
  double g = 0.0, scale = 0.0, anorm = 0.0, s, f, h, c, x, y, z;
  int l = 0, nm = 0;
  for (int i = 0; i < n; i++) {            /* Householder reduction */
    l = i + 1;
    rv1[i] = scale * g;
    g = 0.0; s = 0.0; scale = 0.0;
    for (int k = i; k < n; k++) scale = scale + fabs(UM(k, i));
    if (scale != 0.0) {
      for (int k = i; k < n; k++) {
        UM(k, i) = UM(k, i) / scale;
        s = s + UM(k, i) * UM(k, i);
      }
      f = UM(i, i);
      g = -f_sign(sqrt(s), f);
      h = f * g - s;
      UM(i, i) = f - g;
      for (int j = l; j < n; j++) {
        s = 0.0;
        for (int k = i; k < n; k++) s = s + UM(k, i) * UM(k, j);
        f = s / h;
        for (int k = i; k < n; k++) UM(k, j) = UM(k, j) + f * UM(k, i);
      }
      for (int k = i; k < n; k++) UM(k, i) = scale * UM(k, i);
    }
    W[i] = scale * g;
    g = 0.0; s = 0.0; scale = 0.0;
    for (int k = l; k < n; k++) scale = scale + fabs(UM(i, k));
    if (scale != 0.0) {
      for (int k = l; k < n; k++) {
        UM(i, k) = UM(i, k) / scale;
        s = s + UM(i, k) * UM(i, k);
      }
      f = UM(i, l);
      g = -f_sign(sqrt(s), f);
      h = f * g - s;
      UM(i, l) = f - g;
      for (int k = l; k < n; k++) rv1[k] = UM(i, k) / h;
      for (int j = l; j < n; j++) {
        s = 0.0;
        for (int k = l; k < n; k++) s = s + UM(j, k) * UM(i, k);
        for (int k = l; k < n; k++) UM(j, k) = UM(j, k) + s * rv1[k];
      }
      for (int k = l; k < n; k++) UM(i, k) = scale * UM(i, k);
    }
    anorm = fmax(anorm, fabs(W[i]) + fabs(rv1[i]));
  }
  for (int i = n - 1; i >= 0; i--) {       /* right-hand transformations */
    if (i < n - 1) {
      if (g != 0.0) {
        for (int j = l; j < n; j++) VM(j, i) = (UM(i, j) / UM(i, l)) / g;
        for (int j = l; j < n; j++) {
          s = 0.0;
          for (int k = l; k < n; k++) s = s + UM(i, k) * VM(k, j);
          for (int k = l; k < n; k++) VM(k, j) = VM(k, j) + s * VM(k, i);
        }
      }
      for (int j = l; j < n; j++) { VM(i, j) = 0.0; VM(j, i) = 0.0; }
    }
    VM(i, i) = 1.0;
    g = rv1[i];
    l = i;
  }
  for (int i = n - 1; i >= 0; i--) {       /* left-hand transformations */
    l = i + 1;
    g = W[i];
    for (int j = l; j < n; j++) UM(i, j) = 0.0;
    if (g != 0.0) {
      g = 1.0 / g;
      for (int j = l; j < n; j++) {
        s = 0.0;
        for (int k = l; k < n; k++) s = s + UM(k, i) * UM(k, j);
        f = (s / UM(i, i)) * g;
        for (int k = i; k < n; k++) UM(k, j) = UM(k, j) + f * UM(k, i);
      }
      for (int j = i; j < n; j++) UM(j, i) = UM(j, i) * g;
    } else {
      for (int j = i; j < n; j++) UM(j, i) = 0.0;
    }
    UM(i, i) = UM(i, i) + 1.0;
  }
  for (int k = n - 1; k >= 0; k--) {       /* diagonalise the bidiagonal form */
    for (int its = 1; its <= 30; its++) {
      int flag = 1;
      for (l = k; l >= 0; l--) {
        nm = l - 1;                        /* rv1[0] == 0: the loop ends at l = 0 */
        if (fabs(rv1[l]) + anorm == anorm) { flag = 0; break; }
        if (fabs(W[nm]) + anorm == anorm) break;
      }
      if (flag) {                          /* cancellation of rv1[l], l > 0 */
        c = 0.0; s = 1.0;
        for (int i = l; i <= k; i++) {
          f = s * rv1[i];
          rv1[i] = c * rv1[i];
          if (fabs(f) + anorm == anorm) break;
          g = W[i];
          h = gr_pythag(f, g);
          W[i] = h;
          h = 1.0 / h;
          c = g * h;
          s = -(f * h);
          for (int j = 0; j < n; j++) {
            y = UM(j, nm);
            z = UM(j, i);
            UM(j, nm) = (y * c) + (z * s);
            UM(j, i) = -(y * s) + (z * c);
          }
        }
      }
      z = W[k];
      if (l == k) {                        /* converged */
        if (z < 0.0) {
          W[k] = -z;
          for (int j = 0; j < n; j++) VM(j, k) = -VM(j, k);
        }
        break;
      }
      if (its == 30) return -1;
      x = W[l];                            /* shift from the bottom 2x2 minor */
      nm = k - 1;
      y = W[nm];
      g = rv1[nm];
      h = rv1[k];
      f = ((y - z) * (y + z) + (g - h) * (g + h)) / (2.0 * h * y);
      g = gr_pythag(f, 1.0);
      f = ((x - z) * (x + z) + h * ((y / (f + f_sign(g, f))) - h)) / x;
      c = 1.0; s = 1.0;                    /* next QR transformation */
      for (int j = l; j <= nm; j++) {
        const int i = j + 1;
        g = rv1[i];
        y = W[i];
        h = s * g;
        g = c * g;
        z = gr_pythag(f, h);
        rv1[j] = z;
        c = f / z;
        s = h / z;
        f = (x * c) + (g * s);
        g = -(x * s) + (g * c);
        h = y * s;
        y = y * c;
        for (int jj = 0; jj < n; jj++) {
          x = VM(jj, j);
          z = VM(jj, i);
          VM(jj, j) = (x * c) + (z * s);
          VM(jj, i) = -(x * s) + (z * c);
        }
        z = gr_pythag(f, h);
        W[j] = z;
        if (z != 0.0) {
          z = 1.0 / z;
          c = f * z;
          s = h * z;
        }
        f = (c * g) + (s * y);
        x = -(s * g) + (c * y);
        for (int jj = 0; jj < n; jj++) {
          y = UM(jj, j);
          z = UM(jj, i);
          UM(jj, j) = (y * c) + (z * s);
          UM(jj, i) = -(y * s) + (z * c);
        }
      }
      rv1[l] = 0.0;
      rv1[k] = f;
      W[k] = x;
    }
  }
  return 0;
}

/* x = V diag(1/w) U^T b, terms of zero singular values skipped
 * (svdlib.F90:282-303); b overwritten with x; tmp work (n) */
void or_svd_gr_backsub(int n, const double *U, const double *W, const double *V, double *b,
                       double *tmp) {
  for (int j = 0; j < n; j++) {
    double s = 0.0;
    if (W[j] != 0.0) {
      for (int i = 0; i < n; i++) s = s + UM(i, j) * b[i];
      s = s / W[j];
    }
    tmp[j] = s;
  }
  for (int j = 0; j < n; j++) {
    double s = 0.0;
    for (int jj = 0; jj < n; jj++) s = s + VM(j, jj) * tmp[jj];
    b[j] = s;
  }
}

#undef UM
#undef VM

/* svd_solve (svdlib.F90:22-306) of a column-major n x n A, b overwritten */
int or_svd_solve_gr(int n, const double *A, double *b) {
  double *U = malloc(sizeof(double) * (size_t)n * n * 2 + sizeof(double) * 3 * n);
  if (!U) return -2;
  double *V = U + (size_t)n * n, *W = V + (size_t)n * n, *rv1 = W + n, *tmp = rv1 + n;
  memcpy(U, A, sizeof(double) * (size_t)n * n);
  int rc = or_svd_gr(n, U, W, V, rv1);
  if (rc == 0) or_svd_gr_backsub(n, U, W, V, b, tmp);
  free(U);
  return rc;
}

static int nsize_a(int m) { return (m * m + 3 * m) / 2; }

static double res_tol_of(int m) {
  const int n = nsize_a(m);
  const double c = m == 4 ? 1.0e-3 : m == 6 ? 5.0e-3 : m == 8 ? 3.0e-2 : 1.0;
  return c * (double)(n * n * n * n) * 2.220446049250313080847263336181640625e-16 / (double)m;
}

/* The checks of one trial h (the body of the do-while loops of
 * labf.F90:1151-1336 and :1517-1702 up to the decision): the list src[0..K0)
 * cut to r <= 2 hh (into list, *K entries), the order-m moment system of the
 * first nsize(m) of the filter order's functions, its SVD solves for the
 * Laplacian and the two gradients, the Laplacian residual measure *hcs.
 * amp_all 0: the reference's sequence -- the 33 amplification checks run
 * while hcs < rtol/hh, the first failure setting hcs = 2 rtol/hh.  amp_all 1:
 * *hcs is the residual measure alone and all 33 checks run; *amp_fail is the
 * first failing wavenumber index (1..33) or 0.  Returns 0, or -1 if the SVD
 * did not converge. */
static int eval_hh(or_domain *d, int i, int m, double hh, const int *src, int K0, double *work,
                   int *list, int *K_out, double *hcs_out, int amp_all, int *amp_fail) {
  const int mG = d->forder, nG = nsize_a(mG), n = nsize_a(m);
  const double rtol = res_tol_of(m), amp_tol = 1.0001;
  const int nk = 32;
  double *A = work, *M = A + (size_t)n * n, *U = M + (size_t)n * n, *V = U + (size_t)n * n;
  double *W = V + (size_t)n * n, *rv1 = W + n, *tmp = rv1 + n;
  double *bL = tmp + n, *bX = bL + n, *bY = bX + n, *xv = bY + n;
  double *g = xv + nG, *mono = g + nG;
  /* temporary list: current neighbours within the reduced support */
  int K = 0;
  for (int k = 0; k < K0; k++) {
    const int j = src[k];
    const double rx = d->x[i] - d->x[j], ry = d->y[i] - d->y[j];
    const double rad = sqrt(rx * rx + ry * ry);
    if (rad <= hh * 2.0) list[K++] = j;
  }
  *K_out = K;
  memset(A, 0, sizeof(double) * (size_t)n * n);
  for (int k = 0; k < K; k++) {
    const int j = list[k];
    const double rx = d->x[i] - d->x[j], ry = d->y[i] - d->y[j];
    const double x = -rx, y = -ry;
    const double rad = sqrt(rx * rx + ry * ry);
    const double ff1 = or_wab(rad / hh, d->h2);
    or_abfs(mG, x / hh, y / hh, g);
    for (int a = 0; a < nG; a++) g[a] = g[a] * ff1;
    or_monomials(mG, x / hh, y / hh, mono);
    for (int r = 0; r < n; r++)
      for (int c = 0; c < n; c++) A[(size_t)c * n + r] = A[(size_t)c * n + r] + mono[r] * g[c];
  }
  memcpy(M, A, sizeof(double) * (size_t)n * n);
  /* the three solves share one decomposition: svd_solve recomputes the
   * same one from the same matrix each time */
  memcpy(U, A, sizeof(double) * (size_t)n * n);
  if (or_svd_gr(n, U, W, V, rv1)) return -1;
  for (int a = 0; a < n; a++) { bL[a] = 0.0; bX[a] = 0.0; bY[a] = 0.0; }
  bL[2] = 1.0 / hh / hh;
  bL[4] = 1.0 / hh / hh;
  bX[0] = 1.0 / hh;
  bY[1] = 1.0 / hh;
  or_svd_gr_backsub(n, U, W, V, bL, tmp);
  or_svd_gr_backsub(n, U, W, V, bX, tmp);
  or_svd_gr_backsub(n, U, W, V, bY, tmp);
  /* residual of the Laplacian solution */
  for (int a = 0; a < n; a++) xv[a] = 0.0;
  xv[2] = -1.0 / hh / hh;
  xv[4] = -1.0 / hh / hh;
  double hcs = 0.0;
  for (int r = 0; r < n; r++) {
    for (int c = 0; c < n; c++) xv[r] = xv[r] + M[(size_t)c * n + r] * bL[c];
    hcs = hcs + xv[r] * xv[r];
  }
  hcs = hcs * hh * hh;
  hcs = sqrt(hcs / (double)n);
  const double hcs_raw = hcs;
  if (amp_fail) *amp_fail = 0;
  /* amplification at wavenumbers i1*pi/(s nk), i1 = 1..nk+1 */
  int i1 = 0;
  if (amp_all || hcs < rtol / hh) {
    while (i1 <= nk && (amp_all || hcs < rtol / hh)) {
      i1 = i1 + 1;
      double aL = 0.0, aX = 0.0, aY = 0.0, qq = 0.0;
      for (int k = 0; k < K; k++) {
        const int j = list[k];
        const double rx = d->x[i] - d->x[j], ry = d->y[i] - d->y[j];
        const double x = -rx, y = -ry;
        const double rad = sqrt(rx * rx + ry * ry);
        const double ff1 = or_wab(rad / hh, d->h2);
        or_abfs(mG, x / hh, y / hh, g);
        for (int a = 0; a < nG; a++) g[a] = g[a] * ff1;
        qq = (double)i1 * PI_A / d->s[i] / (double)nk;
        double dL = 0.0, dX = 0.0, dY = 0.0;
        for (int a = 0; a < n; a++) {
          dL = dL + bL[a] * g[a];
          dX = dX + bX[a] * g[a];
          dY = dY + bY[a] * g[a];
        }
        aL = aL + (0.5 - 0.5 * cos(x * qq) * cos(y * qq)) * dL;
        aX = aX + cos(y * qq) * sin(x * qq) * dX;
        aY = aY + cos(x * qq) * sin(y * qq) * dY;
      }
      aL = aL / (qq * qq);
      aX = aX / qq;
      aY = aY / qq;
      int bad = 0;
      if (aL > amp_tol || aL < 0.0) bad = 1;
      if (fabs(aX) > amp_tol) bad = 1;
      if (fabs(aY) > amp_tol) bad = 1;
      if (bad) hcs = 2.0 * rtol / hh;
      if (isnan(hcs)) { hcs = 2.0 * rtol / hh; bad = 1; }
      if (bad && amp_fail && *amp_fail == 0) *amp_fail = i1;
    }
  }
  *hcs_out = amp_all ? hcs_raw : hcs;
  return 0;
}

static size_t adapt_work_size(const or_domain *d) {
  const int nG = nsize_a(d->forder > d->morder ? d->forder : d->morder);
  return (size_t)4 * nG * nG + 12 * (size_t)nG + 16;
}

/* One h-reduction sequence of node i (labf.F90:1151-1336 and :1517-1702).
 * pass 1: the filter order's system on the full list, accepted reductions
 * shorten the list (and set h = h_small); pass 2: the derivative order's
 * system (the first nsize unknowns of the filter order's functions) on the
 * small list, accepted reductions set h_small and the small count.  Returns
 * the number of accepted reductions, or -1 if an SVD did not converge. */
static int adapt_node(or_domain *d, int i, int pass, double *work, int *list) {
  const int m = pass == 1 ? d->forder : d->morder;
  const double reduction = (double)0.98f;
  const double limit = (double)logf(0.6f) / log(reduction);
  const double rtol = res_tol_of(m);
  const size_t base = (size_t)i * d->nplink;
  int ii = 0;
  for (;;) {
    const double hh = (pass == 1 ? d->h[i] : d->hs[i]) * reduction;
    const int K0 = pass == 1 ? d->ij_count[i] : d->ij_count_small[i];
    int K;
    double hcs;
    if (eval_hh(d, i, m, hh, d->ij_link + base, K0, work, list, &K, &hcs, 0, NULL)) return -1;
    if ((double)ii > limit) hcs = 2.0 * rtol / hh;   /* at most ~26 reductions */
    if (hcs >= rtol / hh) break;
    /* accept: the reduced h and list */
    ii = ii + 1;
    if (pass == 1) {
      d->h[i] = hh;
      d->hs[i] = hh;
      for (int k = 0; k < d->nplink; k++) d->ij_link[base + k] = k < K ? list[k] : 0;
      d->ij_count[i] = K;
      d->ij_count_small[i] = K;
    } else {
      d->hs[i] = hh;
      d->ij_count_small[i] = K;
    }
  }
  return ii;
}

/* The reference's acceptance checks of the stencils h_sel (npfb values, h of
 * pass 1 or h_small of pass 2) on the current lists, which must hold every
 * neighbour within 2 h_sel (the lists before adaptation do): per node the
 * residual measure over its tolerance, ratio[i] = hcs / (rtol / h_sel), and
 * amp[i] = the first of the 33 wavenumbers whose amplification check fails
 * (0: all pass).  Nodes that adaptation leaves alone (node_type <= 2, 998)
 * get ratio 0, amp 0.  Returns 0, or -1 if an SVD did not converge. */
int or_adapt_check(or_domain *d, const double *h_sel, int pass, double *ratio, int *amp) {
  const size_t wsz = adapt_work_size(d);
  const int m = pass == 1 ? d->forder : d->morder;
  const double rtol = res_tol_of(m);
  int failed = 0;
#pragma omp parallel for schedule(dynamic, 8) reduction(| : failed)
  for (int i = 0; i < d->npfb; i++) {
    const int nt = d->node_type[i];
    ratio[i] = 0.0;
    amp[i] = 0;
    if (nt <= 2 || nt == 998) continue;
    double *work = malloc(sizeof(double) * wsz);
    int *list = malloc(sizeof(int) * (size_t)d->nplink);
    int K, af = 0;
    double hcs = 0.0;
    if (eval_hh(d, i, m, h_sel[i], d->ij_link + (size_t)i * d->nplink, d->ij_count[i], work,
                list, &K, &hcs, 1, &af))
      failed = 1;
    ratio[i] = hcs / (rtol / h_sel[i]);
    amp[i] = af;
    free(work);
    free(list);
  }
  return failed ? -1 : 0;
}

/* adapt_stencils + adapt_stencils_main over the real nodes (2D: every node is
 * in the first layer), after or_find_neighbours.  Boundary nodes and their
 * first FD rows (node_type <= 2) and order-reduced nodes (998) keep h.
 * reductions (nullable, 2*npfb): accepted reductions per node, pass 1 then
 * pass 2.  Returns 0, or -1 if an SVD did not converge (the reference stops). */
int or_adapt_stencils(or_domain *d, int *reductions) {
  const size_t wsz = adapt_work_size(d);
  int failed = 0;
  for (int pass = 1; pass <= 2; pass++) {
#pragma omp parallel for schedule(dynamic, 8) reduction(| : failed)
    for (int i = 0; i < d->npfb; i++) {
      const int nt = d->node_type[i];
      int r = 0;
      if (nt > 2 && nt != 998) {
        double *work = malloc(sizeof(double) * wsz);
        int *list = malloc(sizeof(int) * (size_t)d->nplink);
        r = adapt_node(d, i, pass, work, list);
        free(work);
        free(list);
        if (r < 0) { failed = 1; r = 0; }
      }
      if (reductions) reductions[(size_t)(pass - 1) * d->npfb + i] = r;
    }
  }
  return failed ? -1 : 0;
}

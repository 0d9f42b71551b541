// k_adapt.hip -- stencil adaptation on the device: adapt_stencils
// (labf.F90:1079-1416) and adapt_stencils_main (:1418-1721).
//
// Each interior node's h is reduced by 2% at a time while (1) the residual of
// the SVD solution of its Laplacian moment system stays below res_tol/hh and
// (2) its d/dx, d/dy and Laplacian operators amplify none of 33 wavenumbers up
// to the node's Nyquist wavenumber by more than amp_tol, for at most ~26
// reductions; pass 1 does this with the filter order on the full stencil (h
// and h_small together), pass 2 with the derivative order on the small
// stencil (h_small).  The residual test compares quantities at the level of
// the SVD's own rounding, so the SVD here is the Golub-Reinsch algorithm in the
// Numerical Recipes formulation svdlib.F90:22-306 uses, with every matrix
// element receiving the same floating-point operations in the same order:
// one wavefront per node, a lane per column (left Householder steps, the
// accumulations) or per row (right Householder steps, the rotations), each
// lane's inner sums sequential, the scalar recurrences evaluated identically
// in every lane.  Contraction into fused multiply-adds is off in this file,
// so the moment matrices, the SVD, the solutions and the residual are the
// oracle's (and the reference's) bit for bit; Wab's (1 - q/2)^4 is the
// correctly rounded fourth power (the reference calls pow(t, 4.0)).  The
// amplification sums use the device's cos/sin and a wave reduction: they are
// compared with 1.0001 and 0, not with rounding-level thresholds.
//
// One block (one wavefront) per real node; LDS holds the node's SVD (U, V,
// the original matrix), a batch of monomial/ABF rows and per-neighbour data.
#pragma clang fp contract(off)

#include <math.h>

#include <algorithm>

#include "sv_internal.h"

namespace {

constexpr int WV = SV_WAVE;
constexpr int nsz(int m) { return (m * m + 3 * m) / 2; }
constexpr int RB = 8;       // monomial/ABF rows per batch of the matrix build

// Probabilists' Hermite polynomials as labf.F90:2329-2391 writes them
__device__ double herm(int n, double z) {
  double z2, z3, z4;
  switch (n) {
  case 0: return 1.0;
  case 1: return z;
  case 2: return z * z - 1.0;
  case 3: return z * z * z - 3.0 * z;
  case 4: return z * z * z * z - 6.0 * z * z + 3.0;
  case 5: z2 = z * z; z3 = z * z2; return z2 * z3 - 10.0 * z3 + 15.0 * z;
  case 6: z2 = z * z; z3 = z * z2; return z3 * z3 - 15.0 * z2 * z2 + 45.0 * z2 - 15.0;
  case 7: z2 = z * z; z3 = z * z2; z4 = z3 * z;
    return z4 * z3 - 21.0 * z3 * z2 + 105.0 * z3 - 105.0 * z;
  default: z2 = z * z; z4 = z2 * z2;
    return z4 * z4 - 28.0 * z4 * z2 + 210.0 * z4 - 420.0 * z2 + 105.0;
  }
}

// Taylor monomials (labf.F90:1815-1920) and Hermite ABFs times ff1
// (labf.F90:2017-2120) of one offset, graded order, MG's nsize entries
template <int MG>
__device__ void mono_abf_rows(double x, double y, double ff1, double *mx, double *gx) {
  double xp[MG + 1], yp[MG + 1], hx[MG + 1], hy[MG + 1];
  xp[0] = 1.0; yp[0] = 1.0; xp[1] = x; yp[1] = y;
#pragma unroll
  for (int k = 2; k <= MG; k++) { xp[k] = xp[k - 1] * x; yp[k] = yp[k - 1] * y; }
#pragma unroll
  for (int k = 0; k <= MG; k++) { hx[k] = herm(k, x); hy[k] = herm(k, y); }
  constexpr double fact[11] = {1, 1, 2, 6, 24, 120, 720, 5040, 40320, 362880, 3628800};
  int idx = 0;
#pragma unroll
  for (int dg = 1; dg <= MG; dg++)
#pragma unroll
    for (int b = 0; b <= dg; b++) {
      const int a = dg - b;
      const double coef = 1.0 / (fact[a] * fact[b]);
      double v;
      if (b == 0) v = (a == 1) ? x : coef * xp[a];
      else if (a == 0) v = (b == 1) ? y : coef * yp[b];
      else if (a == 1 && b == 1) v = x * y;
      else if (a == 1) v = (coef * x) * yp[b];
      else if (b == 1) v = (coef * xp[a]) * y;
      else v = (coef * xp[a]) * yp[b];
      mx[idx] = v;
      const double gv = (b == 0) ? hx[a] : (a == 0) ? hy[b] : hx[a] * hy[b];
      gx[idx] = gv * ff1;
      idx++;
    }
}

// t^4 rounded once: t^2 = p + e exactly, t^4 = p^2 + (2 p e + e^2)
__device__ double pow4_rn(double t) {
  const double p = t * t, e = fma(t, t, -p);
  const double q1 = p * p, q2 = fma(p, p, -q1);
  return q1 + (q2 + (2.0 * p * e + e * e));
}

// Wendland C2 with the global normalisation (rbfs.F90:48-57)
__device__ double wab_rn(double q, double h2) {
  const double t = pow4_rn(1.0 - 0.5 * q);
  if (q <= 2.0 && q > 0.0) return (2.0 * q + 1.0) * (7.0 / (4.0 * SV_PI * h2)) * t;
  return 0.0;
}

__device__ double gr_pythag(double a, double b) {
  const double aa = fabs(a), ab = fabs(b);
  if (aa > ab) {
    const double r = ab / aa;
    return aa * sqrt(1.0 + r * r);
  }
  if (ab == 0.0) return 0.0;
  const double r = aa / ab;
  return ab * sqrt(1.0 + r * r);
}

__device__ double f_sign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

// Golub-Reinsch SVD of the n x n column-major matrix in U (LDS): U, W, V, rv1
// as svdlib.F90 leaves them.  Every lane evaluates the scalar recurrences;
// lane j updates column j (or row j).  Returns false without convergence.
template <int N>
__device__ bool svd_gr(double *U, double *W, double *V, double *rv1) {
  constexpr int n = N;
  const int ln = threadIdx.x;
#define UM(r, c) U[(c) * n + (r)]
#define VM(r, c) V[(c) * n + (r)]
  double g = 0.0, scale = 0.0, anorm = 0.0, s, f, h, c, x, y, z;
  int l = 0, nm = 0;
  for (int i = 0; i < n; i++) {
    l = i + 1;
    const double rv1i = scale * g;
    if (ln == 0) rv1[i] = rv1i;
    g = 0.0; s = 0.0; scale = 0.0;
    for (int k = i; k < n; k++) scale = scale + fabs(UM(k, i));
    __syncthreads();
    if (scale != 0.0) {
      if (ln >= i && ln < n) UM(ln, i) = UM(ln, i) / scale;
      __syncthreads();
      for (int k = i; k < n; k++) s = s + UM(k, i) * UM(k, i);
      f = UM(i, i);
      g = -f_sign(sqrt(s), f);
      h = f * g - s;
      __syncthreads();
      if (ln == 0) UM(i, i) = f - g;
      __syncthreads();
      if (ln >= l && ln < n) {
        const int j = ln;
        double sj = 0.0;
        for (int k = i; k < n; k++) sj = sj + UM(k, i) * UM(k, j);
        const double fj = sj / h;
        for (int k = i; k < n; k++) UM(k, j) = UM(k, j) + fj * UM(k, i);
      }
      __syncthreads();
      if (ln >= i && ln < n) UM(ln, i) = scale * UM(ln, i);
      __syncthreads();
    }
    const double wi = scale * g;
    if (ln == 0) W[i] = wi;
    g = 0.0; s = 0.0; scale = 0.0;
    for (int k = l; k < n; k++) scale = scale + fabs(UM(i, k));
    __syncthreads();
    if (scale != 0.0) {
      if (ln >= l && ln < n) UM(i, ln) = UM(i, ln) / scale;
      __syncthreads();
      for (int k = l; k < n; k++) s = s + UM(i, k) * UM(i, k);
      f = UM(i, l);
      g = -f_sign(sqrt(s), f);
      h = f * g - s;
      __syncthreads();
      if (ln == 0) UM(i, l) = f - g;
      __syncthreads();
      if (ln >= l && ln < n) rv1[ln] = UM(i, ln) / h;
      __syncthreads();
      if (ln >= l && ln < n) {
        const int j = ln;
        double sj = 0.0;
        for (int k = l; k < n; k++) sj = sj + UM(j, k) * UM(i, k);
        for (int k = l; k < n; k++) UM(j, k) = UM(j, k) + sj * rv1[k];
      }
      __syncthreads();
      if (ln >= l && ln < n) UM(i, ln) = scale * UM(i, ln);
      __syncthreads();
    }
    anorm = fmax(anorm, fabs(wi) + fabs(rv1[i]));
    __syncthreads();
  }
  for (int i = n - 1; i >= 0; i--) {           // right-hand transformations
    if (i < n - 1) {
      if (g != 0.0) {
        if (ln >= l && ln < n) VM(ln, i) = (UM(i, ln) / UM(i, l)) / g;
        __syncthreads();
        if (ln >= l && ln < n) {
          const int j = ln;
          double sj = 0.0;
          for (int k = l; k < n; k++) sj = sj + UM(i, k) * VM(k, j);
          for (int k = l; k < n; k++) VM(k, j) = VM(k, j) + sj * VM(k, i);
        }
        __syncthreads();
      }
      if (ln >= l && ln < n) { VM(i, ln) = 0.0; VM(ln, i) = 0.0; }
    }
    if (ln == 0) VM(i, i) = 1.0;
    g = rv1[i];
    l = i;
    __syncthreads();
  }
  for (int i = n - 1; i >= 0; i--) {           // left-hand transformations
    l = i + 1;
    g = W[i];
    if (ln >= l && ln < n) UM(i, ln) = 0.0;
    __syncthreads();
    if (g != 0.0) {
      g = 1.0 / g;
      if (ln >= l && ln < n) {
        const int j = ln;
        double sj = 0.0;
        for (int k = l; k < n; k++) sj = sj + UM(k, i) * UM(k, j);
        const double fj = (sj / UM(i, i)) * g;
        for (int k = i; k < n; k++) UM(k, j) = UM(k, j) + fj * UM(k, i);
      }
      __syncthreads();
      if (ln >= i && ln < n) UM(ln, i) = UM(ln, i) * g;
    } else {
      if (ln >= i && ln < n) UM(ln, i) = 0.0;
    }
    __syncthreads();
    if (ln == 0) UM(i, i) = UM(i, i) + 1.0;
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; k--) {           // the bidiagonal form
    for (int its = 1; its <= 30; its++) {
      bool flag = true;
      for (l = k; l >= 0; l--) {
        nm = l - 1;
        if (fabs(rv1[l]) + anorm == anorm) { flag = false; break; }
        if (fabs(W[nm]) + anorm == anorm) break;
      }
      if (flag) {
        c = 0.0; s = 1.0;
        for (int i = l; i <= k; i++) {
          f = s * rv1[i];
          const double rvi = c * rv1[i];
          __syncthreads();
          if (ln == 0) rv1[i] = rvi;
          if (fabs(f) + anorm == anorm) { __syncthreads(); break; }
          g = W[i];
          h = gr_pythag(f, g);
          __syncthreads();
          if (ln == 0) W[i] = h;
          h = 1.0 / h;
          c = g * h;
          s = -(f * h);
          if (ln < n) {
            const double yv = UM(ln, nm), zv = UM(ln, i);
            UM(ln, nm) = (yv * c) + (zv * s);
            UM(ln, i) = -(yv * s) + (zv * c);
          }
          __syncthreads();
        }
      }
      z = W[k];
      if (l == k) {
        if (z < 0.0) {
          __syncthreads();
          if (ln == 0) W[k] = -z;
          if (ln < n) VM(ln, k) = -VM(ln, k);
        }
        __syncthreads();
        break;
      }
      if (its == 30) return false;
      x = W[l];
      nm = k - 1;
      y = W[nm];
      g = rv1[nm];
      h = rv1[k];
      f = ((y - z) * (y + z) + (g - h) * (g + h)) / (2.0 * h * y);
      g = gr_pythag(f, 1.0);
      f = ((x - z) * (x + z) + h * ((y / (f + f_sign(g, f))) - h)) / x;
      c = 1.0; s = 1.0;
      for (int j = l; j <= nm; j++) {
        const int i = j + 1;
        g = rv1[i];
        y = W[i];
        h = s * g;
        g = c * g;
        z = gr_pythag(f, h);
        __syncthreads();
        if (ln == 0) rv1[j] = z;
        c = f / z;
        s = h / z;
        f = (x * c) + (g * s);
        g = -(x * s) + (g * c);
        h = y * s;
        y = y * c;
        if (ln < n) {
          const double xv = VM(ln, j), zv = VM(ln, i);
          VM(ln, j) = (xv * c) + (zv * s);
          VM(ln, i) = -(xv * s) + (zv * c);
        }
        z = gr_pythag(f, h);
        if (ln == 0) W[j] = z;
        if (z != 0.0) {
          z = 1.0 / z;
          c = f * z;
          s = h * z;
        }
        f = (c * g) + (s * y);
        x = -(s * g) + (c * y);
        if (ln < n) {
          const double yv = UM(ln, j), zv = UM(ln, i);
          UM(ln, j) = (yv * c) + (zv * s);
          UM(ln, i) = -(yv * s) + (zv * c);
        }
        __syncthreads();
      }
      if (ln == 0) {
        rv1[l] = 0.0;
        rv1[k] = f;
        W[k] = x;
      }
      __syncthreads();
    }
  }
  return true;
#undef UM
#undef VM
}

// x = V diag(1/w) U^T b with zero singular values skipped (svdlib.F90:282-303);
// b and tmp in LDS, lane j computes entry j
template <int N>
__device__ void svd_backsub(const double *U, const double *W, const double *V, double *b,
                            double *tmp) {
  const int ln = threadIdx.x;
  if (ln < N) {
    double s = 0.0;
    if (W[ln] != 0.0) {
      for (int i = 0; i < N; i++) s = s + U[ln * N + i] * b[i];
      s = s / W[ln];
    }
    tmp[ln] = s;
  }
  __syncthreads();
  double s = 0.0;
  if (ln < N)
    for (int jj = 0; jj < N; jj++) s = s + V[jj * N + ln] * tmp[jj];
  __syncthreads();
  if (ln < N) b[ln] = s;
  __syncthreads();
}

struct AdaptArgs {
  int npfb;
  const double *x, *y, *s;
  double *h, *hs;
  const int *node_type;
  const int64_t *row_ptr;
  int *col, *cnt, *cnt_small;   // lists, rewritten by pass 1 (reduced stencils)
  double h2;
  double red, limit;            // 0.98 and log(0.6)/log(0.98) as the reference holds them
  double rtol[2];               // res_tol of pass 1 (filter order) and pass 2
  int *reductions;              // [2][npfb]
  int *fail;
  int kcap;                     // longest list
};

// LDS of one node (doubles): U, V, M [NG*NG]; rows X, G [RB][NG]; W, rv1,
// tmp, bL, bX, bY, xv [NG]; per neighbour x, y, dL, dX, dY [kcap]; list ints
template <int MF>
constexpr int lds_doubles_fixed() {
  return 3 * nsz(MF) * nsz(MF) + 2 * RB * nsz(MF) + 7 * nsz(MF);
}

// One pass over node i: returns the accepted reductions, or -1 (no SVD
// convergence).  M = the pass's order (n unknowns), MF the filter order (the
// functions evaluated, NG of them).
template <int MF, int M>
__device__ int adapt_pass(const AdaptArgs &a, int i, int pass, double *lds) {
  constexpr int NG = nsz(MF), n = nsz(M);
  constexpr int NE = (n * n + WV - 1) / WV;
  const int ln = threadIdx.x;
  double *U = lds, *V = U + NG * NG, *Mx = V + NG * NG;
  double *RX = Mx + NG * NG, *RG = RX + RB * NG;
  double *Wv = RG + RB * NG, *rv1 = Wv + NG, *tmp = rv1 + NG, *bL = tmp + NG, *bX = bL + NG,
         *bY = bX + NG, *xv = bY + NG;
  double *kx = xv + NG, *ky = kx + a.kcap, *dL = ky + a.kcap, *dX = dL + a.kcap,
         *dY = dX + a.kcap;
  int *list = reinterpret_cast<int *>(dY + a.kcap);
  const double xi = a.x[i], yi = a.y[i];
  const int64_t base = a.row_ptr[i];
  const double rtol = a.rtol[pass - 1], amp_tol = 1.0001;
  const int nk = 32;
  int ii = 0;
  for (;;) {
    const double hh = (pass == 1 ? a.h[i] : a.hs[i]) * a.red;
    const int K0 = pass == 1 ? a.cnt[i] : a.cnt_small[i];
    // the current list within the reduced support, in list order
    int K = 0;
    for (int k0 = 0; k0 < K0; k0 += WV) {
      const int k = k0 + ln;
      int j = 0;
      bool keep = false;
      if (k < K0) {
        j = a.col[base + k];
        const double rx = xi - a.x[j], ry = yi - a.y[j];
        keep = sqrt(rx * rx + ry * ry) <= hh * 2.0;
      }
      const unsigned long long bal = __ballot(keep);
      const int pos = K + __popcll(bal & ((1ull << ln) - 1ull));
      if (keep) {
        list[pos] = j;
        kx[pos] = -(xi - a.x[j]);
        ky[pos] = -(yi - a.y[j]);
      }
      K += __popcll(bal);
    }
    __syncthreads();
    // moment matrix A(r, c) = sum_k X_k[r] g_k[c], k in list order
    double acc[NE];
#pragma unroll
    for (int m = 0; m < NE; m++) acc[m] = 0.0;
    for (int k0 = 0; k0 < K; k0 += RB) {
      if (ln < RB && k0 + ln < K) {   // a row per lane
        const double x = kx[k0 + ln], y = ky[k0 + ln];
        const double rad = sqrt(x * x + y * y);
        mono_abf_rows<MF>(x / hh, y / hh, wab_rn(rad / hh, a.h2), RX + ln * NG, RG + ln * NG);
      }
      __syncthreads();
      const int nb = min(RB, K - k0);
      for (int b = 0; b < nb; b++) {
#pragma unroll
        for (int m = 0; m < NE; m++) {
          const int e = ln + m * WV;
          if (e < n * n) acc[m] = acc[m] + RX[b * NG + e % n] * RG[b * NG + e / n];
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < NE; m++) {
      const int e = ln + m * WV;
      if (e < n * n) { U[e] = acc[m]; Mx[e] = acc[m]; }
    }
    if (ln < n) {
      bL[ln] = (ln == 2 || ln == 4) ? 1.0 / hh / hh : 0.0;
      bX[ln] = ln == 0 ? 1.0 / hh : 0.0;
      bY[ln] = ln == 1 ? 1.0 / hh : 0.0;
    }
    __syncthreads();
    if (!svd_gr<n>(U, Wv, V, rv1)) return -1;
    svd_backsub<n>(U, Wv, V, bL, tmp);
    svd_backsub<n>(U, Wv, V, bX, tmp);
    svd_backsub<n>(U, Wv, V, bY, tmp);
    // residual of the Laplacian solution: row r by lane r, then the sum of
    // squares in row order
    if (ln < n) {
      double v = (ln == 2 || ln == 4) ? -1.0 / hh / hh : 0.0;
      for (int c = 0; c < n; c++) v = v + Mx[c * n + ln] * bL[c];
      xv[ln] = v;
    }
    __syncthreads();
    double hcs = 0.0;
    for (int r = 0; r < n; r++) hcs = hcs + xv[r] * xv[r];
    hcs = hcs * hh * hh;
    hcs = sqrt(hcs / (double)n);
    if (hcs < rtol / hh) {
      // per neighbour: the three operators' weights (psi . g_k)
      for (int k0 = 0; k0 < K; k0 += WV) {
        const int k = k0 + ln;
        if (k < K) {
          const double x = kx[k], y = ky[k];
          const double rad = sqrt(x * x + y * y);
          double mrow[NG], grow[NG];
          mono_abf_rows<MF>(x / hh, y / hh, wab_rn(rad / hh, a.h2), mrow, grow);
          double sL = 0.0, sX = 0.0, sY = 0.0;
#pragma unroll
          for (int q = 0; q < n; q++) {
            sL = sL + bL[q] * grow[q];
            sX = sX + bX[q] * grow[q];
            sY = sY + bY[q] * grow[q];
          }
          dL[k] = sL;
          dX[k] = sX;
          dY[k] = sY;
        }
      }
      __syncthreads();
      int i1 = 0;
      while (i1 <= nk && hcs < rtol / hh) {
        i1 = i1 + 1;
        const double qq = (double)i1 * SV_PI / a.s[i] / (double)nk;
        double tL = 0.0, tX = 0.0, tY = 0.0;
        for (int k = ln; k < K; k += WV) {
          const double x = kx[k], y = ky[k];
          tL += (0.5 - 0.5 * cos(x * qq) * cos(y * qq)) * dL[k];
          tX += cos(y * qq) * sin(x * qq) * dX[k];
          tY += cos(x * qq) * sin(y * qq) * dY[k];
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
          tL += __shfl_xor(tL, o);
          tX += __shfl_xor(tX, o);
          tY += __shfl_xor(tY, o);
        }
        const double aL = tL / (qq * qq), aX = tX / qq, aY = tY / qq;
        if (aL > amp_tol || aL < 0.0) hcs = 2.0 * rtol / hh;
        if (fabs(aX) > amp_tol) hcs = 2.0 * rtol / hh;
        if (fabs(aY) > amp_tol) hcs = 2.0 * rtol / hh;
        if (isnan(hcs)) hcs = 2.0 * rtol / hh;
      }
    }
    if ((double)ii > a.limit) hcs = 2.0 * rtol / hh;
    if (hcs >= rtol / hh) break;
    ii = ii + 1;
    if (pass == 1) {
      for (int k = ln; k < K; k += WV) a.col[base + k] = list[k];
      if (ln == 0) {
        a.h[i] = hh;
        a.hs[i] = hh;
        a.cnt[i] = K;
        a.cnt_small[i] = K;
      }
    } else if (ln == 0) {
      a.hs[i] = hh;
      a.cnt_small[i] = K;
    }
    __threadfence_block();
    __syncthreads();
  }
  return ii;
}

template <int MF, int MD>
__global__ __launch_bounds__(WV) void k_adapt(AdaptArgs a) {
  extern __shared__ double lds[];
  const int i = blockIdx.x;
  if (i >= a.npfb) return;
  const int nt = a.node_type[i];
  int r1 = 0, r2 = 0;
  if (nt > 2 && nt != 998) {
    r1 = adapt_pass<MF, MF>(a, i, 1, lds);
    if (r1 >= 0) r2 = adapt_pass<MF, MD>(a, i, 2, lds);
    if (r1 < 0 || r2 < 0) {
      if (threadIdx.x == 0) atomicAdd(a.fail, 1);
      r1 = max(r1, 0);
      r2 = max(r2, 0);
    }
  }
  if (threadIdx.x == 0) {
    a.reductions[i] = r1;
    a.reductions[a.npfb + i] = r2;
  }
}

template <int MF, int MD>
int launch_adapt(sv_ctx *c, const AdaptArgs &a) {
  const size_t shm = sizeof(double) * ((size_t)lds_doubles_fixed<MF>() + 5 * (size_t)a.kcap) +
                     sizeof(int) * (size_t)a.kcap + 64;
  if (shm > 64 * 1024)
    return sv_fail(c, SV_E_ARG, "adapt_stencils: %d-entry lists need %zu bytes of LDS (> 64 KiB)",
                   a.kcap, shm);
  k_adapt<MF, MD><<<a.npfb, WV, shm, c->stream>>>(a);
  SV_HIP(c, hipGetLastError());
  return 0;
}

double res_tol_of(int m) {
  const int n = nsz(m);
  const double cf = m == 4 ? 1.0e-3 : m == 6 ? 5.0e-3 : 3.0e-2;
  return cf * (double)(n * n * n * n) * 2.220446049250313080847263336181640625e-16 / (double)m;
}

}  // namespace

// adapt_stencils on the device lists (after sv_find_neighbours): the node h,
// h_small and the lists' counts change; the caller rebuilds the lists.
int sv_dev_adapt(sv_ctx *c, int *reductions_host) {
  const int n = c->npfb;
  const int mf = c->cfg.forder, md = c->cfg.morder;
  if (md > mf)
    return sv_fail(c, SV_E_ARG, "adapt_stencils needs morder <= forder (%d > %d): the "
                                "derivative system takes the filter order's first functions",
                   md, mf);
  if (mf > 8)
    return sv_fail(c, SV_E_ARG, "adapt_stencils: LABFM order %d is not supported on the device "
                                "(its SVD takes a lane per column, at most 64)", mf);
  DevTmp<int> red, fail;
  SV_HIP(c, red.alloc(2 * (size_t)n));
  SV_HIP(c, fail.alloc(1));
  SV_HIP(c, hipMemsetAsync(fail, 0, sizeof(int), c->stream));
  AdaptArgs a{};
  a.npfb = n;
  a.x = c->x; a.y = c->y; a.s = c->s; a.h = c->h; a.hs = c->hs;
  a.node_type = c->node_type;
  a.row_ptr = c->row_ptr;
  a.col = c->col_raw; a.cnt = c->cnt; a.cnt_small = c->cnt_small;
  a.h2 = c->h2;
  a.red = (double)0.98f;                              // labf.F90:1129, single literal
  a.limit = (double)logf(0.6f) / log((double)0.98f);  // labf.F90:1306
  a.rtol[0] = res_tol_of(mf);
  a.rtol[1] = res_tol_of(md);
  a.reductions = red;
  a.fail = fail;
  a.kcap = std::max(c->maxK, 1);
  int rc;
  if (mf == 4) rc = launch_adapt<4, 4>(c, a);
  else if (mf == 6) rc = md == 4 ? launch_adapt<6, 4>(c, a) : launch_adapt<6, 6>(c, a);
  else rc = md == 4 ? launch_adapt<8, 4>(c, a) : md == 6 ? launch_adapt<8, 6>(c, a)
                                                           : launch_adapt<8, 8>(c, a);
  if (rc) return rc;
  int nfail = 0;
  SV_HIP(c, hipMemcpyAsync(&nfail, fail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  if (reductions_host)
    SV_HIP(c, hipMemcpyAsync(reductions_host, red, sizeof(int) * 2 * (size_t)n,
                             hipMemcpyDeviceToHost, c->stream));
  SV_HIP(c, hipStreamSynchronize(c->stream));
  if (nfail)
    return sv_fail(c, SV_E_SINGULAR, "adapt_stencils: SVD not converging on %d nodes "
                                     "(svdlib.F90:218-221 stops)", nfail);
  return 0;
}

// k_bound.hip -- boundary nodes: normals and the boundary operators of
// calc_boundary_weights (labf.F90:505-1039).
//
// A boundary node (type 0 wall, 1 inflow, 2 outflow) is followed in node order
// by its four FD rows (types -1..-4) at s, 2s, 3s, 4s along its unit normal
// (setup_domain.F90:179-193).  Its operators: 5-point one-sided differences
// along the normal, and along the tangent a 1D LABFM of order 6 -- Hermite
// ABFs times the Wendland C2 derivative kernel (rbfs.F90:58-68), a 6x6 moment
// system over the same-layer neighbours.  Rows 1 and 2 get 5-point normal
// differences and a 1D LABFM along a tangent turned by the tangential
// gradient of s.  The counts are small (one system per boundary node and row),
// so each system is one thread: LU with partial pivoting and one refinement
// step on a 6x6 in registers.  Three launches keep the reference's order of
// dependencies: rows read their parent's tangential weights before the outflow
// reversal.
#include "sv_internal.h"

namespace {

constexpr int BT = 64;

// rbfs.F90:58-68: the Wendland C2 derivative kernel (7/(4 pi h0^3)) times
// (5/8 q^4 - 15/4 q^3 + 15/2 q^2 - 5 q) on 0 < q <= 2, in Horner form
__device__ __forceinline__ double fac_b(double qq, double h3) {
  if (!(qq <= 2.0 && qq > 0.0)) return 0.0;
  const double poly = qq * (-5.0 + qq * (7.5 + qq * (-3.75 + qq * 0.625)));
  return poly * (7.0 / (4.0 * SV_PI * h3));
}

// monomials1D (labf.F90:1041-1048): z^k / k!, k = 1..6, by the recurrence
// t_k = t_{k-1} z / k
__device__ __forceinline__ void mono1(double z, double c[6]) {
  double t = z;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    c[k] = t;
    t = t * z / (double)(k + 2);
  }
}

// abfs1D (labf.F90:1050-1077): Hermite ABFs He_1..He_6 (probabilists'), from
// the three-term recurrence He_{k+1} = z He_k - k He_{k-1}, times ff1; the
// reference scales He_3 by 1/sqrt(2) (:1060)
__device__ __forceinline__ void abf1(double z, double ff1, double c[6]) {
  double hm = 1.0, hk = z;   // He_0, He_1
#pragma unroll
  for (int k = 1; k <= 6; k++) {
    c[k - 1] = ff1 * hk;
    const double hn = z * hk - (double)k * hm;
    hm = hk;
    hk = hn;
  }
  c[2] *= 0.70710678118654752440;
}

// 6x6 solve A x = b for NR right-hand sides (A column-major, A[c*6+r]):
// Gaussian elimination with partial pivoting, then one step of iterative
// refinement with the residual of the original matrix.
template <int NR>
__device__ __forceinline__ void solve6(const double A0[36], double x[NR][6]) {
  double A[36];
  int piv[6];
#pragma unroll
  for (int e = 0; e < 36; e++) A[e] = A0[e];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    int p = k;
    double v = fabs(A[k * 6 + k]);
    for (int r = k + 1; r < 6; r++)
      if (fabs(A[k * 6 + r]) > v) { v = fabs(A[k * 6 + r]); p = r; }
    piv[k] = p;
    // the row swap through compile-time indexes only (a run-time index into
    // A would place it in scratch memory)
#pragma unroll
    for (int r = k + 1; r < 6; r++)
      if (r == p)
#pragma unroll
        for (int c = 0; c < 6; c++) { const double t = A[c * 6 + k]; A[c * 6 + k] = A[c * 6 + r]; A[c * 6 + r] = t; }
    const double d = A[k * 6 + k] == 0.0 ? 1.0 : A[k * 6 + k];
    for (int r = k + 1; r < 6; r++) {
      const double m = A[k * 6 + r] / d;
      A[k * 6 + r] = m;
      for (int c = k + 1; c < 6; c++) A[c * 6 + r] -= m * A[c * 6 + k];
    }
  }
  auto lu_solve = [&](double y[6]) {
#pragma unroll
    for (int k = 0; k < 6; k++)
#pragma unroll
      for (int r = k + 1; r < 6; r++)
      {
        const bool sw = piv[k] == r;
        const double yk = y[k], yr = y[r];
        y[k] = sw ? yr : yk;
        y[r] = sw ? yk : yr;
      }
    for (int r = 1; r < 6; r++)
      for (int c = 0; c < r; c++) y[r] -= A[c * 6 + r] * y[c];
    for (int r = 5; r >= 0; r--) {
      for (int c = r + 1; c < 6; c++) y[r] -= A[c * 6 + r] * y[c];
      const double d = A[r * 6 + r] == 0.0 ? 1.0 : A[r * 6 + r];
      y[r] = y[r] / d;
    }
  };
#pragma unroll
  for (int q = 0; q < NR; q++) {
    double b[6], y[6];
    for (int r = 0; r < 6; r++) { b[r] = x[q][r]; y[r] = b[r]; }
    lu_solve(y);
    double res[6];
    for (int r = 0; r < 6; r++) {
      double s = b[r];
      for (int c = 0; c < 6; c++) s -= A0[c * 6 + r] * y[c];
      res[r] = s;
    }
    lu_solve(res);
    for (int r = 0; r < 6; r++) x[q][r] = y[r] + res[r];
  }
}

struct BndArgs {
  int npfb, nb;
  const int *blist;
  const int *nt;              // node types [np]
  const double *x, *y, *s, *h, *rnx, *rny;
  const int64_t *row_ptr;
  int *col;                   // ghost ids (col_raw)
  int *cnt, *cnt_small;
  double *wx, *wy, *wl, *wh, *vol;
  double *ooR;                // [nb]
  double *wtt;                // [nb][wtt_ld] tangential second-derivative weights
  int wtt_ld;
  double h3;
};

// FD parent of real node j (rows only), -1 otherwise
__device__ __forceinline__ int fd_parent(const BndArgs &a, int j) {
  if (j >= a.npfb) return -1;
  const int t = a.nt[j];
  return t < 0 ? j + t : -1;
}

// Tangential coordinate of j relative to i (labf.F90:704-712)
__device__ __forceinline__ double tcoord(const BndArgs &a, int i, int j, double xt, double yt) {
  const double rx = a.x[i] - a.x[j], ry = a.y[i] - a.y[j];
  double t = -(rx * xt + ry * yt);
  if (i == j) t = 0.0;
  return t;
}

// 1D moment matrix of node i over its same-layer neighbours (:695-717)
__device__ __forceinline__ void assemble1d(const BndArgs &a, int i, int64_t base, int K, double xt, double yt,
                           double hh, double A[36]) {
  for (int e = 0; e < 36; e++) A[e] = 0.0;
  const int ti = a.nt[i];
  for (int k = 0; k < K; k++) {
    const int j = a.col[base + k];
    if (a.nt[j] != ti) continue;
    double t = tcoord(a, i, j, xt, yt);
    double xv[6], g[6];
    mono1(t / hh, xv);
    const double ff1 = fac_b(fabs(t / hh), a.h3);
    t = t / hh;
    abf1(t, ff1, g);
    for (int r = 0; r < 6; r++)
      for (int c = 0; c < 6; c++) A[c * 6 + r] += xv[r] * g[c];
  }
}

__device__ __forceinline__ double w1d(const BndArgs &a, int i, int j, double xt, double yt,
                                      double hh, const double b[6]) {
  double t = tcoord(a, i, j, xt, yt);
  const double ff1 = fac_b(fabs(t / hh), a.h3);
  t = t / hh;
  double g[6];
  abf1(t, ff1, g);
  double s = 0.0;
  for (int q = 0; q < 6; q++) s += b[q] * g[q];
  return s;
}

// Boundary node: neighbour pruning (:541-566), normal differences (:568-612),
// tangential 1D LABFM (:692-758), wall curvature (:929-945) and the wall
// Laplacian (:970-980); node volume s^2 (:490-494).
__global__ __launch_bounds__(BT) void k_bnd_nodes(BndArgs a) {
  const int jb = blockIdx.x * BT + threadIdx.x;
  if (jb >= a.nb) return;
  const int i = a.blist[jb], ti = a.nt[i];
  const int64_t base = a.row_ptr[i];
  const int n0 = a.cnt[i];
  int K = 0;
  for (int k = 0; k < n0; k++) {
    const int j = a.col[base + k];
    const bool keep = (a.nt[j] == ti) || (fd_parent(a, j) == i);
    if (keep) a.col[base + K++] = j;
  }
  for (int k = 0; k < n0; k++) {
    a.wx[base + k] = 0.0; a.wy[base + k] = 0.0; a.wl[base + k] = 0.0; a.wh[base + k] = 0.0;
  }
  a.cnt[i] = K;
  a.cnt_small[i] = K;
  const double dx = a.s[i], dx2 = dx * dx;
  for (int k = 0; k < K; k++) {
    const int j = a.col[base + k];
    if (fd_parent(a, j) != i) continue;
    const int tj = a.nt[j];
    a.wx[base + k] = tj == -1 ? 4.0 / dx : tj == -2 ? -3.0 / dx : tj == -3 ? (4.0 / 3.0) / dx
                                                                            : -0.25 / dx;
  }
  const double xt = -a.rny[i], yt = a.rnx[i], hh = a.h[i];
  double *wt = a.wtt + (size_t)jb * a.wtt_ld;
  for (int k = 0; k < a.wtt_ld; k++) wt[k] = 0.0;
  double A[36];
  assemble1d(a, i, base, K, xt, yt, hh, A);
  double b[3][6];
  for (int q = 0; q < 3; q++)
    for (int r = 0; r < 6; r++) b[q][r] = 0.0;
  b[0][0] = 1.0;     // d/dt
  b[1][1] = 1.0;     // d2/dt2
  b[2][3] = -1.0;    // -d4/dt4 (filter)
  solve6<3>(A, b);
  double tt = 0.0;
  for (int k = 0; k < K; k++) {
    const int j = a.col[base + k];
    if (a.nt[j] != ti) continue;
    a.wy[base + k] = w1d(a, i, j, xt, yt, hh, b[0]) / hh;
    const double g2t = w1d(a, i, j, xt, yt, hh, b[1]) / hh / hh;
    wt[k] = g2t;                // ij_wb_grad2(2,k,jj) (:755)
    a.wl[base + k] = g2t;       // the tangential second derivative, for now
    a.wh[base + k] = 0.0 + w1d(a, i, j, xt, yt, hh, b[2]);
    if (ti == 0) tt = tt + g2t * ((a.rnx[j] * a.rnx[i] + a.rny[j] * a.rny[i]) - 1.0);
  }
  const double ooR = ti == 0 ? sqrt(fabs(tt)) : 0.0;
  a.ooR[jb] = ooR;
  for (int k = 0; k < K; k++) {
    if (ti != 0) { a.wl[base + k] = 0.0; continue; }
    const int j = a.col[base + k];
    double g2n = 0.0;
    if (fd_parent(a, j) == i) {
      const int tj = a.nt[j];
      g2n = tj == -1 ? -104.0 / 12.0 / dx2 : tj == -2 ? 114.0 / 12.0 / dx2
          : tj == -3 ? -56.0 / 12.0 / dx2 : 11.0 / 12.0 / dx2;
    }
    g2n = g2n + a.wx[base + k] * ooR;
    a.wl[base + k] = g2n + a.wl[base + k];
  }
  a.vol[i] = a.s[i] * a.s[i];
}

// Rows 1 and 2 (:615-677, :761-923), the wall divergence term of their
// Laplacian (:957-966) and the map of their gradients to x-y (:984-1000).
__global__ __launch_bounds__(BT) void k_bnd_rows(BndArgs a) {
  const int t = blockIdx.x * BT + threadIdx.x;
  if (t >= 2 * a.nb) return;
  const int jb = t >> 1, row = 1 + (t & 1);
  const int par = a.blist[jb], i = par + row;
  const int64_t base = a.row_ptr[i], pb = a.row_ptr[par];
  const int K = a.cnt[i], Kp = a.cnt[par];
  double grads = 0.0;
  for (int k = 0; k < Kp; k++) {
    const int j = a.col[pb + k];
    grads = grads + (a.s[j] - a.s[par]) * a.wy[pb + k];
  }
  grads = -atan(grads);
  const double dx = a.s[i], dx2 = dx * dx;
  for (int k = 0; k < K; k++) { a.wx[base + k] = 0.0; a.wy[base + k] = 0.0; a.wl[base + k] = 0.0; }
  for (int k = 0; k < K; k++) {
    const int j = a.col[base + k];
    if (j >= a.npfb) continue;
    const bool sib = fd_parent(a, j) == par;
    const int tj = a.nt[j];
    double g = 0.0, l = 0.0;
    bool set = false;
    if (row == 1) {
      if (j == par) { g = -0.25 / dx; l = 11.0 / 12.0 / dx2; set = true; }
      if (j == i) { g = 0.0; l = 0.0; set = true; }
      if (sib && tj == -2) { g = 1.5 / dx; l = 6.0 / 12.0 / dx2; set = true; }
      if (sib && tj == -3) { g = -0.5 / dx; l = 4.0 / 12.0 / dx2; set = true; }
      if (sib && tj == -4) { g = 1.0 / 12.0 / dx; l = -1.0 / 12.0 / dx2; set = true; }
    } else {
      if (j == par) { g = 1.0 / 12.0 / dx; l = -1.0 / 12.0 / dx2; set = true; }
      if (sib && tj == -1) { g = -8.0 / 12.0 / dx; l = 16.0 / 12.0 / dx2; set = true; }
      if (j == i) { g = 0.0 / 12.0 / dx; l = 0.0 / 12.0 / dx2; set = true; }
      if (sib && tj == -3) { g = 8.0 / 12.0 / dx; l = 16.0 / 12.0 / dx2; set = true; }
      if (sib && tj == -4) { g = -1.0 / 12.0 / dx; l = -1.0 / 12.0 / dx2; set = true; }
    }
    if (set) { a.wx[base + k] = g; a.wl[base + k] = l; }
  }
  const double xt = cos(grads) * (-a.rny[i]) - sin(grads) * a.rnx[i];
  const double yt = sin(grads) * (-a.rny[i]) + cos(grads) * a.rnx[i];
  const double hh = a.h[i];
  double A[36];
  assemble1d(a, i, base, K, xt, yt, hh, A);
  double b[2][6];
  for (int q = 0; q < 2; q++)
    for (int r = 0; r < 6; r++) b[q][r] = 0.0;
  b[0][0] = 1.0;
  b[1][1] = 1.0;
  solve6<2>(A, b);
  const int ti = a.nt[i];
  for (int k = 0; k < K; k++) {
    const int j = a.col[base + k];
    if (a.nt[j] != ti) continue;
    a.wy[base + k] = w1d(a, i, j, xt, yt, hh, b[0]) / hh;
    a.wl[base + k] = a.wl[base + k] + w1d(a, i, j, xt, yt, hh, b[1]) / hh / hh;
  }
  const double ic = 1.0 / cos(grads), tg = tan(grads);
  for (int k = 0; k < K; k++) a.wy[base + k] = ic * a.wy[base + k] - tg * a.wx[base + k];
  if (a.nt[par] == 0) {
    const double ooR = a.ooR[jb];
    const double den = row == 1 ? (1.0 + ooR * a.s[i]) : (1.0 + 2.0 * ooR * a.s[i]);
    for (int k = 0; k < K; k++) a.wl[base + k] = a.wl[base + k] + a.wx[base + k] * ooR / den;
  }
  const double xn = a.rnx[par], yn = a.rny[par];
  for (int k = 0; k < K; k++) {
    const double gn = a.wx[base + k], gt = a.wy[base + k];
    a.wx[base + k] = xn * gn + (-yn) * gt;
    a.wy[base + k] = yn * gn + xn * gt;
  }
}

// Outflow boundary nodes: their normal-tangent weights reversed into the x-y
// frame (:949-954)
__global__ __launch_bounds__(BT) void k_bnd_finish(BndArgs a) {
  const int jb = blockIdx.x * BT + threadIdx.x;
  if (jb >= a.nb) return;
  const int i = a.blist[jb];
  if (a.nt[i] != 2) return;
  const int64_t base = a.row_ptr[i];
  for (int k = 0; k < a.cnt[i]; k++) {
    a.wx[base + k] = -a.wx[base + k];
    a.wy[base + k] = -a.wy[base + k];
  }
}

__global__ void k_ghost_normals(int npfb, int np, const int *irel, double *rnx, double *rny) {
  const int k = npfb + blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  rnx[k] = rnx[irel[k]];
  rny[k] = rny[irel[k]];
}

}  // namespace

// Normals of the real nodes (rows take their boundary node's, as the IPART
// reader copies them, setup_domain.F90:189) and of the ghosts (their
// parent's, mirror_boundaries.F90:253).
int sv_dev_boundary_normals(sv_ctx *c, const double *rnx_host, const double *rny_host) {
  const int n = c->npfb;
  std::vector<double> nx(rnx_host, rnx_host + n), ny(rny_host, rny_host + n);
  for (int i = 0; i < n; i++)
    if (c->hnt[i] < 0) { nx[i] = nx[i + c->hnt[i]]; ny[i] = ny[i + c->hnt[i]]; }
  for (int jb = 0; jb < c->nb; jb++) {
    const int i = c->hblist[jb];
    const double r = sqrt(nx[i] * nx[i] + ny[i] * ny[i]);
    if (!(fabs(r - 1.0) < 1e-6))
      return sv_fail(c, SV_E_ARG, "boundary node %d: normal (%g, %g) is not a unit vector", i,
                     nx[i], ny[i]);
  }
  if (!c->rnx) {
    if (sv_alloc(c, (void **)&c->rnx, sizeof(double) * c->cap) ||
        sv_alloc(c, (void **)&c->rny, sizeof(double) * c->cap))
      return SV_E_NOMEM;
  }
  SV_HIP(c, hipMemcpy(c->rnx, nx.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  SV_HIP(c, hipMemcpy(c->rny, ny.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  const int ng = c->np - n;
  if (ng > 0) {
    k_ghost_normals<<<(ng + 255) / 256, 256, 0, c->stream>>>(n, c->np, c->irel, c->rnx, c->rny);
    SV_HIP(c, hipGetLastError());
  }
  SV_HIP(c, hipStreamSynchronize(c->stream));
  c->have_normals = true;
  return 0;
}

namespace {
// calc_labf_sums' boundary block (labf.F90:428-443) for the tangential
// second derivative: the row sum, then both divided by L_char^2
__global__ __launch_bounds__(BT) void k_bnd_tt_sums(int nb, const int *blist, const int *cnt,
                                                    double *wtt, int ld, double L,
                                                    double *wtt_sum) {
  const int jb = blockIdx.x * BT + threadIdx.x;
  if (jb >= nb) return;
  const int K = cnt[blist[jb]];
  double *wt = wtt + (size_t)jb * ld;
  double s = 0.0;
  for (int k = 0; k < K; k++) s = s + wt[k];
  wtt_sum[jb] = s / L / L;
  for (int k = 0; k < K; k++) wt[k] = wt[k] / L / L;
}

// The boundary branch of calc_laplacian_transverse_only_on_bound
// (derivatives.F90:327-337): sum over k <= ij_count_small of phi_j w_k, in
// order, minus phi_i times the row sum
__global__ __launch_bounds__(BT) void k_bnd_tt_apply(int nb, const int *blist,
                                                     const int *cnt_small,
                                                     const int64_t *row_ptr, const int *col,
                                                     const int *irel, const double *wtt, int ld,
                                                     const double *wtt_sum, const double *phi,
                                                     double *out) {
  const int jb = blockIdx.x * BT + threadIdx.x;
  if (jb >= nb) return;
  const int i = blist[jb];
  const int K = cnt_small[i];
  const int64_t base = row_ptr[i];
  const double *wt = wtt + (size_t)jb * ld;
  double s = 0.0;
  for (int k = 0; k < K; k++) {
    const int j = irel ? irel[col[base + k]] : col[base + k];
    s = s + phi[j] * wt[k];
  }
  out[i] = s - phi[i] * wtt_sum[jb];
}
}  // namespace

int sv_dev_boundary_tt_sums(sv_ctx *c) {
  if (c->nb == 0 || !c->wtt) return 0;
  k_bnd_tt_sums<<<(c->nb + BT - 1) / BT, BT, 0, c->stream>>>(c->nb, c->blist, c->cnt, c->wtt,
                                                              c->wtt_ld, c->L_char, c->wtt_sum);
  SV_HIP(c, hipGetLastError());
  return 0;
}

int sv_dev_boundary_tt_apply(sv_ctx *c, const double *phi, double *out, bool raw_cols,
                             hipStream_t st) {
  if (c->nb == 0) return 0;
  if (!(c->have_bweights && c->wtt))
    return sv_fail(c, SV_E_ARG, "calc_laplacian_transverse_only_on_bound: "
                   "sv_calc_boundary_weights must come first");
  k_bnd_tt_apply<<<(c->nb + BT - 1) / BT, BT, 0, st>>>(
      c->nb, c->blist, c->cnt_small, c->row_ptr, c->col_raw, raw_cols ? nullptr : c->irel,
      c->wtt, c->wtt_ld, c->wtt_sum, phi, out);
  SV_HIP(c, hipGetLastError());
  return 0;
}

int sv_dev_boundary_weights(sv_ctx *c) {
  if (c->nb == 0) { c->have_bweights = true; return 0; }
  if (!c->ooR && sv_alloc(c, (void **)&c->ooR, sizeof(double) * c->nb)) return SV_E_NOMEM;
  if (c->wtt && c->wtt_ld < c->maxK) {
    sv_free(c, c->wtt);
    c->wtt = nullptr;
  }
  if (!c->wtt) {
    c->wtt_ld = c->maxK > 0 ? c->maxK : 1;
    if (sv_alloc(c, (void **)&c->wtt, sizeof(double) * (size_t)c->nb * c->wtt_ld))
      return SV_E_NOMEM;
  }
  if (!c->wtt_sum && sv_alloc(c, (void **)&c->wtt_sum, sizeof(double) * c->nb))
    return SV_E_NOMEM;
  BndArgs a{};
  a.npfb = c->npfb; a.nb = c->nb; a.blist = c->blist; a.nt = c->node_type;
  a.x = c->x; a.y = c->y; a.s = c->s; a.h = c->h; a.rnx = c->rnx; a.rny = c->rny;
  a.row_ptr = c->row_ptr; a.col = c->col_raw; a.cnt = c->cnt; a.cnt_small = c->cnt_small;
  a.wx = c->wx; a.wy = c->wy; a.wl = c->wl; a.wh = c->wh; a.vol = c->vol; a.ooR = c->ooR;
  a.wtt = c->wtt; a.wtt_ld = c->wtt_ld;
  a.h3 = c->h2 * c->h0;
  const int g1 = (c->nb + BT - 1) / BT, g2 = (2 * c->nb + BT - 1) / BT;
  k_bnd_nodes<<<g1, BT, 0, c->stream>>>(a);
  k_bnd_rows<<<g2, BT, 0, c->stream>>>(a);
  k_bnd_finish<<<g1, BT, 0, c->stream>>>(a);
  SV_HIP(c, hipGetLastError());
  SV_HIP(c, hipStreamSynchronize(c->stream));
  c->have_bweights = true;
  c->sell_dirty = true;
  c->L.valid = false;
  return 0;
}

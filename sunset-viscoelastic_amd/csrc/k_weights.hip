// k_weights.hip -- LABFM weight solves (calc_labf_weights, labf.F90:34-231;
// calc_filter_weights :233-394; calc_node_volumes :453-503; calc_labf_sums
// :396-451; filter_coefficients :1723-1811).
//
// Systems are solved in units of G per wavefront (4 of order 4, 2 of order 6,
// 1 of order 8: one matrix row per lane).  For node i with LABFM order m (n =
// (m^2+3m)/2 unknowns) and stencil length K:
//   1. the moment matrix A = sum_k X_k g_k^T (X_k: Taylor monomials of the
//      scaled offset, g_k: Hermite ABFs times the Wendland C2 kernel) is
//      contracted on the FP64 matrix cores, v_mfma_f64_16x16x4_f64, with
//      16- or 32-neighbour chunks of X and g staged in LDS;
//   2. A x = b for the four right-hand sides (d/dx, d/dy, Laplacian,
//      hyperviscous filter) by Gauss-Jordan elimination with partial pivoting,
//      each row and its right-hand sides in one lane's registers (pivot search
//      by DPP reductions, pivot row broadcast through LDS): N steps in
//      sequence and no substitutions;
//   3. one step of iterative refinement: the residual accumulated in
//      double-double and the elimination replayed on it (N more steps), which
//      brings the solution to the exact solution of the FP64 matrix (the
//      reference's Golub-Reinsch SVD, svdlib.F90, is only within cond(A)*eps
//      of it);
//   4. weights w_k = (psi . g_k)/hh are evaluated one neighbour per lane and
//      written in CSR order.
// The derivative system (h_small, small stencil) and the filter system (h, full
// stencil) are one system when they coincide (same order, h_small == h, same
// stencil), and then share the factorisation (4 right-hand sides).
#include <algorithm>

#include "sv_internal.h"

typedef double dbl4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int nsize(int m) { return (m * m + 3 * m) / 2; }
constexpr int npad(int m) { return ((nsize(m) + 15) / 16) * 16; }
// k! for k <= 10 as a table: a recursive constexpr function used outside a
// constant expression is not folded, and the device code then calls it (a
// call per monomial, with its stack frame and the register budget of a call)
constexpr double kFact[11] = {1.0, 1.0, 2.0, 6.0, 24.0, 120.0, 720.0, 5040.0, 40320.0,
                              362880.0, 3628800.0};

// Probabilists' Hermite polynomials He_0..He_M, written as labf.F90:2329-2391
template <int M>
__device__ __forceinline__ void hermite_all(double z, double *H) {
  H[0] = 1.0;
  H[1] = z;
  if (M >= 2) H[2] = z * z - 1.0;
  if (M >= 3) H[3] = z * z * z - 3.0 * z;
  if (M >= 4) H[4] = z * z * z * z - 6.0 * z * z + 3.0;
  if (M >= 5) {
    const double z2 = z * z, z3 = z * z2;
    H[5] = z2 * z3 - 10.0 * z3 + 15.0 * z;
    if (M >= 6) H[6] = z3 * z3 - 15.0 * z2 * z2 + 45.0 * z2 - 15.0;
    if (M >= 7) {
      const double z4 = z3 * z;
      H[7] = z4 * z3 - 21.0 * z3 * z2 + 105.0 * z3 - 105.0 * z;
    }
    if (M >= 8) {
      const double z4 = z2 * z2;
      H[8] = z4 * z4 - 28.0 * z4 * z2 + 210.0 * z4 - 420.0 * z2 + 105.0;
    }
    if (M >= 9) {
      const double z4 = z3 * z;
      H[9] = z4 * z3 * z2 - 36.0 * z4 * z3 + 378.0 * z3 * z2 - 1260.0 * z3 + 945.0 * z;
    }
    if (M >= 10) {
      const double z4 = z2 * z2;
      H[10] = z4 * z4 * z2 - 45.0 * z4 * z4 + 630.0 * z4 * z2 - 3150.0 * z4 + 4725.0 * z2 - 945.0;
    }
  }
}

// Wendland C2, rbfs.F90:48-57
__device__ __forceinline__ double wab(double q, double h2) {
  double t = 1.0 - 0.5 * q;
  t = t * t;
  t = t * t;
  return (q <= 2.0 && q > 0.0) ? (2.0 * q + 1.0) * (7.0 / (4.0 * SV_PI * h2)) * t : 0.0;
}

// Row of Taylor monomials x^a y^b/(a!b!) in graded order (labf.F90:1815-1920)
template <int M, int STRIDE = 1>
__device__ __forceinline__ void mono_row(double x, double y, double *dst) {
  double xp[M + 1], yp[M + 1];
  xp[0] = 1.0; yp[0] = 1.0; xp[1] = x; yp[1] = y;
#pragma unroll
  for (int k = 2; k <= M; k++) { xp[k] = xp[k - 1] * x; yp[k] = yp[k - 1] * y; }
  int idx = 0;
#pragma unroll
  for (int d = 1; d <= M; d++)
#pragma unroll
    for (int b = 0; b <= d; b++) {
      const int a = d - b;
      const double coef = 1.0 / (kFact[a] * kFact[b]);
      double v;
      if (b == 0) v = (a == 1) ? x : coef * xp[a];
      else if (a == 0) v = (b == 1) ? y : coef * yp[b];
      else if (a == 1 && b == 1) v = x * y;
      else if (a == 1) v = (coef * x) * yp[b];
      else if (b == 1) v = (coef * xp[a]) * y;
      else v = (coef * xp[a]) * yp[b];
      dst[STRIDE * idx++] = v;
    }
}

// Row of Hermite ABFs He_a(x) He_b(y) * w (labf.F90:2017-2120)
template <int M, int STRIDE = 1>
__device__ __forceinline__ void abf_row(double x, double y, double w, double *dst) {
  double hx[M + 1], hy[M + 1];
  hermite_all<M>(x, hx);
  hermite_all<M>(y, hy);
  int idx = 0;
#pragma unroll
  for (int d = 1; d <= M; d++)
#pragma unroll
    for (int b = 0; b <= d; b++) {
      const int a = d - b;
      const double v = (b == 0) ? hx[a] : (a == 0) ? hy[b] : hx[a] * hy[b];
      dst[STRIDE * idx++] = v * w;
    }
}

// psi . g for one right-hand side, g evaluated on the fly: sum over the
// unknowns in order of psi[a] * (He_p(x) He_q(y) w)
template <int M>
__device__ __forceinline__ double abf_dot(double x, double y, double w, const double *psi) {
  double hx[M + 1], hy[M + 1];
  hermite_all<M>(x, hx);
  hermite_all<M>(y, hy);
  double acc = 0.0;
  int idx = 0;
#pragma unroll
  for (int d = 1; d <= M; d++)
#pragma unroll
    for (int b = 0; b <= d; b++) {
      const int a = d - b;
      const double v = ((b == 0) ? hx[a] : (a == 0) ? hy[b] : hx[a] * hy[b]) * w;
      acc += psi[idx] * v;
      idx++;
    }
  return acc;
}

__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Hyperviscous filter right-hand side, labf.F90:325-334 (0-based indices)
__device__ __forceinline__ double hyp_rhs(int m, int r) {
  if (m <= 5) return r == 9 ? -1.0 : r == 11 ? -2.0 : r == 13 ? -1.0 : 0.0;
  if (m <= 7) return (r == 20 || r == 26) ? 1.0 : (r == 22 || r == 24) ? 3.0 : 0.0;
  if (m <= 9)
    return (r == 35 || r == 43) ? -1.0 : (r == 37 || r == 41) ? -4.0 : r == 39 ? -6.0 : 0.0;
  if (m <= 11)
    return (r == 54 || r == 64) ? 1.0 : (r == 56 || r == 62) ? 5.0
           : (r == 58 || r == 60) ? 10.0 : 0.0;
  return 0.0;
}

enum { SYS_D = 1, SYS_F = 2 };

struct SysArgs {
  const int *sys_node;       // node per system
  const int *sys_kind;       // SYS_D | SYS_F bits
  int nsys;
  const double *x, *y, *h, *hs;
  const int *node_type;
  const int64_t *row_ptr;
  const int *col, *cnt, *cnt_small;
  double h2;
  double *wx, *wy, *wl, *wh, *vol;
  int *singular;
  int *queue;                // persistent launches: next unit to take
  const int *sys_order;      // mixed-order launch: LABFM order of each system
  const int *unit_first;     // mixed-order launch: first system of each unit
  int nunits;                //   (a unit: up to G_m systems of one order m)
};

// Packing of systems into a wavefront.  The factorisation and the
// substitutions hold one matrix row per lane, so a wavefront takes G = 64 / S
// systems of the same order, each on a group of S lanes (the smallest of 16,
// 32, 64 that holds the N rows): 4 order-4, 2 order-6 or 1 order-8 system.
template <int M>
struct Pack {
  static constexpr int N = nsize(M);
  static constexpr int S = N <= 16 ? 16 : N <= 32 ? 32 : 64;
  static constexpr int G = SV_WAVE / S;
  // register budget: waves per SIMD (order 8 holds two 44-double rows)
  static constexpr int W = M <= 6 ? 2 : 1;
};

// LDS of one wavefront (doubles), one area reused by the phases:
//   assembly     the CH-neighbour chunks of monomial and ABF rows ([2][CH][NP]),
//                then the assembled matrix ([N][LDA], row-major; LDA odd, so
//                lanes reading different rows hit different banks), which its
//                group's lanes take into registers before the next assembly
//   elimination  two pivot-row buffers ([2][G][UBS]: the row's entries past the
//                pivot, 1 / pivot at N, its four right-hand sides at N+1..N+4),
//                used by alternate steps;
//                after them Xs [G][NP][4]: the solution for the residual, then
//                the refinement's per-step values, then the refined solution
template <int M, int CH_>
struct Lds {
  static constexpr int N = nsize(M), NP = npad(M), LDA = (N % 2 == 1) ? N : N + 1;
  static constexpr int CH = CH_;
  static constexpr int G = Pack<M>::G;
  static constexpr int ASZ = N * LDA;
  // chunks are stored column-major (function f of neighbour k at f * LDC + k,
  // LDC = CH + 1 odd: the lanes writing one function hit distinct banks;
  // row-major, all rows started on the same bank).  Row-major 32-neighbour
  // chunks measured slower than 16-neighbour ones; column-major they tie.
  static constexpr bool TR = true;
  static constexpr int LDC = TR ? CH + 1 : 1;
  static constexpr int CHUNK = TR ? 2 * LDC * NP : 2 * CH * NP;
  static constexpr int UBS = NP + 8;
  static constexpr int XOFF = 2 * G * UBS;
  static constexpr int XS = XOFF + G * NP * 4;
  static constexpr int T1 = CHUNK > ASZ ? CHUNK : ASZ;
  static constexpr int TOTAL = T1 > XS ? T1 : XS;
};

// Moment matrix A = sum_k X_k g_k^T (labf.F90:88-123) on the FP64 matrix
// cores: 16- or 32-neighbour chunks of Taylor monomial rows X and ABF rows g
// are staged in LDS (Xc, Gc) and contracted with
// v_mfma_f64_16x16x4_f64; the result is written to Adst ([NP][LDA], LDS).
template <int M, int CH_>
__device__ __forceinline__ void assemble(const SysArgs &S_, int i, double hh, int Kuse,
                                         int64_t base, double *Xc, double *Gc, double *Adst) {
  using L = Lds<M, CH_>;
  constexpr int N = L::N, NP = L::NP, LDA = L::LDA, CH = L::CH, T = NP / 16;
  constexpr bool TR = L::TR;
  constexpr int LDC = L::LDC;
  const int l = threadIdx.x;
  const double xi = S_.x[i], yi = S_.y[i];
  dbl4 acc[T][T];
#pragma unroll
  for (int a = 0; a < T; a++)
#pragma unroll
    for (int b = 0; b < T; b++) acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
  const int role = l / CH;   // 0: monomial row, 1: ABF row, 2-3 (CH = 16): idle
  // scaled offsets by one reciprocal (the weights move by rounding only: ~1
  // ulp in the moment matrix, far inside the 1e-10 weight tolerance)
  const double rhh = 1.0 / hh;
  for (int c0 = 0; c0 < Kuse; c0 += CH) {
    const int kk = c0 + (l % CH);
    double *row = (role == 0 ? Xc : Gc) + (TR ? (l % CH) : (l % CH) * NP);
    if (role < 2 && kk < Kuse) {
      const int j = S_.col[base + kk];
      const double x = -(xi - S_.x[j]), y = -(yi - S_.y[j]);
      const double xs = x * rhh, ys = y * rhh;
      if (role == 0) {
        mono_row<M, TR ? LDC : 1>(xs, ys, row);
      } else {
        const double rad = sqrt(x * x + y * y) * rhh;
        abf_row<M, TR ? LDC : 1>(xs, ys, wab(rad, S_.h2), row);
      }
      // the padding functions N..NP-1 reach only the rows and columns of A
      // that are dropped below, and are left unset
    } else if (role < 2) {
#pragma unroll
      for (int a = 0; a < N; a++) row[TR ? a * LDC : a] = 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < CH / 4; ks++) {
      const int kr = ks * 4 + (l >> 4);
#pragma unroll
      for (int ti = 0; ti < T; ti++) {
        const int fa = ti * 16 + (l & 15);
        const double av = TR ? Xc[fa * LDC + kr] : Xc[kr * NP + fa];
#pragma unroll
        for (int tj = 0; tj < T; tj++) {
          const int fb = tj * 16 + (l & 15);
          const double bv = TR ? Gc[fb * LDC + kr] : Gc[kr * NP + fb];
          acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[ti][tj], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  // D layout of a 16x16 f64 tile (v_mfma_f64_16x16x4_f64): lane l holds rows
  // (l>>4) + 4v, v = 0..3, of column l&15; the padding rows and columns
  // (>= N) are dropped.
#pragma unroll
  for (int ti = 0; ti < T; ti++)
#pragma unroll
    for (int tj = 0; tj < T; tj++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = ti * 16 + (l >> 4) + 4 * v, c = tj * 16 + (l & 15);
        if (r < N && c < N) Adst[r * LDA + c] = acc[ti][tj][v];
      }
  __syncthreads();
}

// One lane's double from lane (l xor 1 / 2 ...) pattern CTRL within its row of
// 16 (DPP; every source lane is inside the row)
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// One lane's double from the partner row (S = 32: rows 0 <-> 1, 2 <-> 3) or
// half (S = 64: lanes l <-> l ^ 32) at the same position, by
// v_permlane16_swap / v_permlane32_swap with both operands v: the first
// result holds the partner's value in the odd rows (upper half), the second
// in the even rows (lower half)
template <int S>
__device__ __forceinline__ double swap_d(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  if (S == 32) {
    const auto rl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto rh = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    const bool odd = (threadIdx.x >> 4) & 1;
    return __hiloint2double(odd ? rh[0] : rh[1], odd ? rl[0] : rl[1]);
  }
  const auto rl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  const auto rh = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  const bool up = threadIdx.x >= 32;
  return __hiloint2double(up ? rh[0] : rh[1], up ? rl[0] : rl[1]);
}

// Maximum of v over each group of S consecutive lanes, in every lane: four
// DPP steps inside each row of 16 (quad_perm [1,0,3,2], [2,3,0,1],
// row_half_mirror, row_mirror), then register swaps across rows for S = 32
// and across halves for S = 64 (no trip through the LDS unit)
template <int S>
__device__ __forceinline__ double group_max(double v) {
  v = fmax(v, dpp_d<0xB1>(v));
  v = fmax(v, dpp_d<0x4E>(v));
  v = fmax(v, dpp_d<0x141>(v));
  v = fmax(v, dpp_d<0x140>(v));
  if (S >= 32) v = fmax(v, swap_d<32>(v));
  if (S >= 64) v = fmax(v, swap_d<64>(v));
  return v;
}

// Right-hand sides of row r (q0 d/dx, q1 d/dy, q2 Laplacian e3+e5, q3 filter)
__device__ __forceinline__ void rhs_row(int r, bool sysD, bool sysF, int mf, double b[4]) {
  b[0] = (sysD && r == 0) ? 1.0 : 0.0;
  b[1] = (sysD && r == 1) ? 1.0 : 0.0;
  b[2] = (sysD && (r == 2 || r == 4)) ? 1.0 : 0.0;
  b[3] = sysF ? hyp_rhs(mf, r) : 0.0;
}

// The system of group g of a unit (G consecutive entries of the sorted list)
struct SysOf {
  int i, kind, Kuse, Kfull, nkeep, mf;
  bool sysD, sysF;
  double hh;
  int64_t base;
};

template <int M>
__device__ __forceinline__ SysOf sys_of(const SysArgs &S_, int sys) {
  constexpr int N = nsize(M);
  SysOf o;
  o.i = S_.sys_node[sys];
  o.kind = S_.sys_kind[sys];
  o.sysD = o.kind & SYS_D;
  o.sysF = o.kind & SYS_F;
  o.hh = o.sysD ? S_.hs[o.i] : S_.h[o.i];
  o.Kuse = o.sysD ? S_.cnt_small[o.i] : S_.cnt[o.i];
  o.Kfull = S_.cnt[o.i];
  o.base = S_.row_ptr[o.i];
  // order reduction near walls by identity padding (labf.F90:126-148, :311-350)
  const int nt = S_.node_type[o.i];
  o.nkeep = N;
  o.mf = M;
  if (M >= 7 && (nt == 998 || nt < 0)) { o.nkeep = 27; o.mf = 6; }
  if (M >= 5 && o.sysD && (nt == -1 || nt == -2)) o.nkeep = 14;
  if (M >= 5 && o.sysF && nt == -1) { o.nkeep = 14; o.mf = 4; }
  return o;
}

// Weights w_k = (psi_q . g_k)/hh (labf.F90:183-215, k <= ij_count) and the
// node volume (calc_node_volumes, labf.F90:453-503) of system `sys`.  Two
// lanes per neighbour, 32 neighbours per pass: lane p = l & 1 keeps psi of
// right-hand sides 2p and 2p+1 (p = 0: d/dx, d/dy; p = 1: Laplacian, filter)
// in registers and evaluates the neighbour's ABFs once for both sums, each in
// the order of the unknowns.
template <int M>
__device__ __forceinline__ void eval_weights(const SysArgs &S_, int sys, const double *Xs) {
  constexpr int N = nsize(M);
  const int l = threadIdx.x, p = l & 1, kl = l >> 1;
  const SysOf o = sys_of<M>(S_, sys);
  const int K = o.Kfull;
  const bool wantB = p == 0 ? o.sysD : o.sysF;   // d/dy or the filter
  const double hh = o.hh, rhh = 1.0 / hh;
  const double xi = S_.x[o.i], yi = S_.y[o.i];
  double psA[N], psB[N];
#pragma unroll
  for (int a = 0; a < N; a++) {
    psA[a] = Xs[a * 4 + 2 * p];
    psB[a] = Xs[a * 4 + 2 * p + 1];
  }
  double *const dA = p ? S_.wl : S_.wx;
  double *const dB = p ? S_.wh : S_.wy;
  double vsum = 0.0;
  for (int k = kl; k < K; k += SV_WAVE / 2) {
    const int j = S_.col[o.base + k];
    const double x = -(xi - S_.x[j]), y = -(yi - S_.y[j]);
    const double rad = sqrt(x * x + y * y) * rhh;
    const double w = wab(rad, S_.h2);
    double hx[M + 1], hy[M + 1];
    hermite_all<M>(x * rhh, hx);
    hermite_all<M>(y * rhh, hy);
    double accA = 0.0, accB = 0.0;
    int idx = 0;
#pragma unroll
    for (int d = 1; d <= M; d++)
#pragma unroll
      for (int b = 0; b <= d; b++) {
        const int a = d - b;
        const double v = ((b == 0) ? hx[a] : (a == 0) ? hy[b] : hx[a] * hy[b]) * w;
        accA += psA[idx] * v;
        accB += psB[idx] * v;
        idx++;
      }
    if (o.sysD) dA[o.base + k] = p ? accA * rhh * rhh : accA * rhh;
    if (wantB) dB[o.base + k] = p ? accB : accB * rhh;
    if (p == 1 && o.sysF) vsum += (9.0 / SV_PI) * exp(-9.0 * rad * rad) / (hh * hh);
  }
  if (o.sysF) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) vsum += __shfl_xor(vsum, off);
    if (l == 0) S_.vol[o.i] = 1.0 / vsum;
  }
}

// Four lanes per neighbour (orders 4 and 8, where two cost a wave per SIMD in
// registers: order 4 measured 6% slower with two): weights
// w_k = (psi_q . g_k)/hh (labf.F90:183-215) and the
// node volume (calc_node_volumes, labf.F90:453-503) of the `count` (<= G)
// systems of a unit.  Four lanes per neighbour, lane q = l & 3 for
// right-hand side q (d/dx, d/dy, Laplacian, filter), 16 neighbours per pass
// of each system; each lane keeps its column of every system's psi in
// registers, and one pass evaluates a neighbour of each system (independent
// chains that the compiler interleaves).  A lane past a system's stencil
// evaluates the node itself and stores nothing.
template <int M, int G>
__device__ __forceinline__ void eval_weights4(const SysArgs &S_, int first, int count,
                                              const double *Xs) {
  constexpr int N = nsize(M), NP = npad(M);
  const int l = threadIdx.x, q = l & 3, kl = l >> 2;
  int ii[G], K[G];
  int64_t bs[G];
  bool want[G], sf[G];
  double hh[G], rhh[G], xi[G], yi[G], vsum[G];
  double *dst[G];
  double ps[G][N];
  int kmax = 0;
#pragma unroll
  for (int gg = 0; gg < G; gg++) {
    const bool on = gg < count;
    const SysOf o = sys_of<M>(S_, first + (on ? gg : 0));
    ii[gg] = o.i;
    K[gg] = on ? o.Kfull : 0;
    bs[gg] = o.base;
    want[gg] = on && (q < 3 ? o.sysD : o.sysF);
    sf[gg] = on && o.sysF;
    hh[gg] = o.hh;
    rhh[gg] = 1.0 / o.hh;
    xi[gg] = S_.x[o.i];
    yi[gg] = S_.y[o.i];
    vsum[gg] = 0.0;
    // (a two-level select: a four-way one can be lowered to a table in scratch)
    double *const d01 = (q & 1) ? S_.wy : S_.wx;
    double *const d23 = (q & 1) ? S_.wh : S_.wl;
    dst[gg] = (q & 2) ? d23 : d01;
#pragma unroll
    for (int a = 0; a < N; a++) ps[gg][a] = Xs[gg * NP * 4 + a * 4 + q];
    kmax = K[gg] > kmax ? K[gg] : kmax;
  }
  for (int k = kl; k < kmax; k += SV_WAVE / 4) {
#pragma unroll
    for (int gg = 0; gg < G; gg++) {
      const bool in = k < K[gg];
      // col has one int past the last row, so the clamped read stays inside it
      const int jr = S_.col[bs[gg] + (in ? k : 0)];
      const int j = in ? jr : ii[gg];
      const double x = -(xi[gg] - S_.x[j]), y = -(yi[gg] - S_.y[j]);
      const double rad = sqrt(x * x + y * y) * rhh[gg];
      const double o = abf_dot<M>(x * rhh[gg], y * rhh[gg], wab(rad, S_.h2), ps[gg]);
      if (in && want[gg])
        dst[gg][bs[gg] + k] = q < 2 ? o * rhh[gg] : q == 2 ? o * rhh[gg] * rhh[gg] : o;
      if (in && q == 3 && sf[gg])
        vsum[gg] += (9.0 / SV_PI) * exp(-9.0 * rad * rad) / (hh[gg] * hh[gg]);
    }
  }
#pragma unroll
  for (int gg = 0; gg < G; gg++) {
    if (sf[gg]) {
      double v = vsum[gg];
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
      if (l == 0) S_.vol[ii[gg]] = 1.0 / v;
    }
  }
}

// Solve the `count` (<= G) systems first .. first+count-1 of one order; the
// block is one wavefront.
//  1. each moment matrix A = sum_k X_k g_k^T (labf.F90:88-123) is contracted on
//     the matrix cores into LDS (assemble), rows and columns beyond nkeep set
//     to the identity, and lane r of group g takes row r of system g into
//     registers (twice: the row to eliminate and the original for the
//     residual);
//  2. Gauss-Jordan elimination with partial pivoting, rows staying in their
//     lanes and the four right-hand sides (d/dx, d/dy, Laplacian, filter)
//     carried along: at step k the group's largest |a[k]| among the rows not
//     yet chosen (lowest row on ties) is the pivot, its row and right-hand
//     sides are broadcast through LDS, and every other row stores its
//     multiplier in a[k] and updates a[k+1..N-1] and its right-hand sides; the
//     pivot row of step k then holds pivot_k x_k;
//  3. one refinement step: the residual b - A x accumulated in double-double,
//     the elimination's multipliers replayed on it (one broadcast per step),
//     and the correction added (the reference's Golub-Reinsch SVD,
//     svdlib.F90, is only within cond(A)*eps of the exact solution of the FP64
//     matrix; the refined solution is at it to rounding);
//  4. the weights of the unit's systems, two (order 6) or four lanes per
//     neighbour (eval_weights, eval_weights4).
template <int M, int CH_>
__device__ __forceinline__ void solve_unit(const SysArgs &S_, int first, int count, double *lds) {
  using P = Pack<M>;
  using L = Lds<M, CH_>;
  constexpr int N = P::N, S = P::S, LDA = L::LDA, NP = L::NP, UBS = L::UBS;
  double *Xs = lds + L::XOFF;           // [G][NP][4]
  const int l = threadIdx.x, g = l / S, r = l % S;
  const bool active = g < count;
  const bool row = active && r < N;

  // ---- 1. assembly, one system after the other ----
  double a[N], a0[N];
#pragma unroll
  for (int c = 0; c < N; c++) { a[c] = 0.0; a0[c] = 0.0; }
  for (int gg = 0; gg < count; gg++) {
    const SysOf o = sys_of<M>(S_, first + gg);
    assemble<M, CH_>(S_, o.i, o.hh, o.Kuse, o.base, lds, lds + L::CHUNK / 2, lds);
    if (o.nkeep < N && l < N) {
      // a rolled loop: unrolled, its N identity entries would be hoisted to
      // the kernel entry and held in registers throughout
#pragma unroll 1
      for (int rr = 0; rr < N; rr++)
        if (rr >= o.nkeep || l >= o.nkeep) lds[rr * LDA + l] = (rr == l) ? 1.0 : 0.0;
    }
    __syncthreads();
    if (row && g == gg) {
#pragma unroll
      for (int c = 0; c < N; c++) {
        a[c] = lds[r * LDA + c];
        a0[c] = a[c];
        if (c % 8 == 7) __builtin_amdgcn_sched_barrier(0);
      }
    }
    __syncthreads();   // the next assembly reuses the area
  }
  const SysOf o = sys_of<M>(S_, first + (active ? g : 0));

  // ---- 2. Gauss-Jordan elimination with partial pivoting ----
  double b[4];
  rhs_row(r, o.sysD, o.sysF, o.mf, b);
  int ps = row ? -1 : N;     // step at which this row was the pivot (N: padding lane)
  double rinv = 1.0;         // 1 / pivot of this row
  bool sing = false;
  const int gbase = g * S;
  // the group maximum of the next step's candidates; each step starts the
  // next step's search as soon as its column k+1 is updated, so the DPP chain
  // overlaps the rest of the row update
  auto candidate = [&](int k) {
    double c = ps < 0 ? fabs(a[k]) : -1.0;
    if (c != c) c = 0.0;   // a NaN entry still yields a pivot (flagged below)
    return c;
  };
  double mx = group_max<S>(candidate(0));
#pragma unroll
  for (int k = 0; k < N; k++) {
    const bool open = ps < 0;
    const double cand = candidate(k);
    const unsigned long long bal = __ballot(open && cand == mx);
    unsigned long long mg = bal >> gbase;
    if (S < 64) mg &= (1ull << S) - 1ull;
    const int p = mg ? __builtin_ctzll(mg) : S;
    const bool piv = open && r == p;
    // step k's buffer; step k+1 uses the other one, and step k+2 writes this
    // one again only after step k+1's barrier, which every read of step k
    // precedes
    double *u = lds + (k & 1) * (P::G * UBS) + g * UBS;
    if (piv) {
      ps = k;
      sing = !(a[k] != 0.0);   // zero or NaN
      rinv = 1.0 / (a[k] == 0.0 ? 1.0 : a[k]);
#pragma unroll
      for (int j = k + 1; j < N; j++) u[j] = a[j];
      u[N] = rinv;
#pragma unroll
      for (int q = 0; q < 4; q++) u[N + 1 + q] = b[q];
    }
    __syncthreads();
    // every row but the pivot takes multiplier lk = a[k] / pivot; the pivot
    // row (and a padding lane) runs the same instructions with lk = 0, which
    // leaves its entries as they are (up to the sign of an exact zero), so
    // the update is one basic block with the next step's search
    const bool upd = row && !piv;
    const double lk = upd ? a[k] * u[N] : 0.0;
    a[k] = upd ? lk : a[k];
    if (k + 1 < N) {
      a[k + 1] = fma(-lk, u[k + 1], a[k + 1]);
      mx = group_max<S>(candidate(k + 1));
    }
#pragma unroll
    for (int j = k + 2; j < N; j++) a[j] = fma(-lk, u[j], a[j]);
#pragma unroll
    for (int q = 0; q < 4; q++) b[q] = fma(-lk, u[N + 1 + q], b[q]);
  }

  // ---- 3. one refinement step ----
  double *xg = Xs + g * NP * 4;
  double xk[4];
#pragma unroll
  for (int q = 0; q < 4; q++) xk[q] = b[q] * rinv;
  if (ps >= 0 && ps < N) {
    double2 *d = reinterpret_cast<double2 *>(xg + ps * 4);
    d[0] = double2{xk[0], xk[1]};
    d[1] = double2{xk[2], xk[3]};
  }
  __syncthreads();
  // residual of row r: b_r - sum_c A0[r][c] x_c in double-double
  double res[4];
  {
    double b0[4];
    rhs_row(r, o.sysD, o.sysF, o.mf, b0);
    double sh[4] = {0.0, 0.0, 0.0, 0.0}, sl[4] = {0.0, 0.0, 0.0, 0.0};
    if (row) {
#pragma unroll
      for (int c = 0; c < N; c++) {
        const double av = a0[c];
        const double2 *d = reinterpret_cast<const double2 *>(xg + c * 4);
        const double2 x01 = d[0], x23 = d[1];
        const double xv[4] = {x01.x, x01.y, x23.x, x23.y};
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const double pr = av * xv[q];
          const double pe = fma(av, xv[q], -pr);
          double t;
          two_sum(sh[q], pr, sh[q], t);
          sl[q] += t + pe;
        }
        if (c % 4 == 3) __builtin_amdgcn_sched_barrier(0);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      double d, de;
      two_sum(b0[q], -sh[q], d, de);
      res[q] = d + (de - sl[q]);
    }
  }
  __syncthreads();   // every residual read of xg is done
  // the elimination replayed on the residual: step k's pivot row gives its
  // value (slot k), every other row subtracts its multiplier times it
#pragma unroll
  for (int k = 0; k < N; k++) {
    if (ps == k) {
      double2 *d = reinterpret_cast<double2 *>(xg + k * 4);
      d[0] = double2{res[0], res[1]};
      d[1] = double2{res[2], res[3]};
    }
    __syncthreads();
    if (ps != k && ps < N) {
      const double2 *d = reinterpret_cast<const double2 *>(xg + k * 4);
      const double2 y01 = d[0], y23 = d[1];
      res[0] = fma(-a[k], y01.x, res[0]);
      res[1] = fma(-a[k], y01.y, res[1]);
      res[2] = fma(-a[k], y23.x, res[2]);
      res[3] = fma(-a[k], y23.y, res[3]);
    }
  }
  __syncthreads();   // every read of the slots is done
  if (ps >= 0 && ps < N) {
    double2 *d = reinterpret_cast<double2 *>(xg + ps * 4);
    d[0] = double2{xk[0] + res[0] * rinv, xk[1] + res[1] * rinv};
    d[1] = double2{xk[2] + res[2] * rinv, xk[3] + res[3] * rinv};
  }
  if (sing && row) atomicAdd(S_.singular, 1);
  __syncthreads();

  // ---- 4. weights, one system after the other ----
  for (int e0 = 0; e0 < count; e0++) {
    if (M == 6) eval_weights<M>(S_, first + e0, Xs + e0 * NP * 4);
    else eval_weights4<M, 1>(S_, first + e0, 1, Xs + e0 * NP * 4);
  }
  __syncthreads();
}

// Systems [u*G, u*G + G) of this launch's list per block
template <int M, int CH_>
__global__ __launch_bounds__(SV_WAVE, Pack<M>::W) void k_weights(SysArgs a) {
  extern __shared__ double lds[];
  constexpr int G = Pack<M>::G;
  const int first = blockIdx.x * G;
  if (first >= a.nsys) return;
  solve_unit<M, CH_>(a, first, min(G, a.nsys - first), lds);
}

// Persistent variant: a fixed grid takes units of G systems from a queue;
// systems are sorted by (order, stencil length) so every wave works on the
// longest remaining systems first.
// (one wave per SIMD: with the unit loop around the solve, the two-wave
// register budget of k_weights would spill)
template <int M, int CH_>
__global__ __launch_bounds__(SV_WAVE, 1) void k_weights_persistent(SysArgs a) {
  extern __shared__ double lds[];
  __shared__ int next;
  constexpr int G = Pack<M>::G;
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(a.queue, 1);
    __syncthreads();
    const int first = next * G;
    __syncthreads();
    if (first >= a.nsys) break;
    solve_unit<M, CH_>(a, first, min(G, a.nsys - first), lds);
  }
}

// One persistent launch over the systems of every order: the queue holds
// units (each G_m systems of one order m, unit_first[u] its first system) in
// the order 8 first, then 6, then 4, each by decreasing stencil length, so the
// grid drains the most expensive systems first and the cheap ones fill the
// tail.  LDS is sized for the largest need of the orders present.
template <int CH_>
__global__ __launch_bounds__(SV_WAVE, 1) void k_weights_mixed(SysArgs a) {
  extern __shared__ double lds[];
  __shared__ int next;
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(a.queue, 1);
    __syncthreads();
    const int u = next;
    __syncthreads();
    if (u >= a.nunits) break;
    const int first = a.unit_first[u];
    const int count = (u + 1 < a.nunits ? a.unit_first[u + 1] : a.nsys) - first;
    const int m = a.sys_order[first];
    if (m == 8) solve_unit<8, CH_>(a, first, count, lds);
    else if (m == 6) solve_unit<6, CH_>(a, first, count, lds);
    else solve_unit<4, CH_>(a, first, count, lds);
  }
}

// ---- systems of more than 64 unknowns (order 10: n = 65) ----
// One system per block of BW = 320 threads (five wavefronts); thread t < 4n
// owns row r = t / 4 of the moment matrix and its columns c = q + 4m (q =
// t % 4).  The entries are assembled in registers from 16-neighbour chunks of
// monomial / ABF rows staged in LDS (a plain FP64 contraction: 65 rows do not
// fit the one-row-per-lane layout of k_weights), copied to LDS for an LU with
// partial pivoting and row swaps, and kept in registers for the refinement
// residual, which each row's four threads sum in double-double and combine
// with two lane exchanges.  Substitutions run one step per barrier with a row
// per thread group; one refinement step; weights one neighbour per thread.
template <int M>
struct Big {
  static constexpr int N = nsize(M), LDA = N + 1, MC = (N + 3) / 4, CH = 16, BW = 320;
  // A [N][LDA], chunks X, G [CH][N], solution / right-hand sides [4][N],
  // residuals [4][N], the permutation [N] and a few shared scalars
  static constexpr int TOTAL = N * LDA + 2 * CH * N + 8 * N + N + 8;
};

template <int M>
__global__ __launch_bounds__(320, 1) void k_weights_big(SysArgs S_) {
  using B = Big<M>;
  constexpr int N = B::N, LDA = B::LDA, MC = B::MC, CH = B::CH, BW = B::BW;
  extern __shared__ double lds[];
  double *A = lds;                       // [N][LDA]
  double *Xc = A + N * LDA;              // [CH][N]
  double *Gc = Xc + CH * N;              // [CH][N]
  double *Xv = Gc + CH * N;              // [4][N]
  double *Rv = Xv + 4 * N;               // [4][N]
  int *perm = reinterpret_cast<int *>(Rv + 4 * N);
  double *shd = Rv + 4 * N + N;          // shared scalars
  const int t = threadIdx.x, r = t >> 2, q = t & 3;
  const bool own = t < 4 * N;
  const int sys = blockIdx.x;
  if (sys >= S_.nsys) return;
  const SysOf o = sys_of<M>(S_, sys);
  const double xi = S_.x[o.i], yi = S_.y[o.i], rhh = 1.0 / o.hh;

  // ---- assembly: A(r, c) = sum_k X_k[r] g_k[c] ----
  double acc[MC];
#pragma unroll
  for (int m = 0; m < MC; m++) acc[m] = 0.0;
  const int role = t / CH;   // 0: monomial row, 1: ABF row, others idle
  for (int c0 = 0; c0 < o.Kuse; c0 += CH) {
    const int kk = c0 + (t % CH);
    if (role < 2) {
      double *row = (role == 0 ? Xc : Gc) + (t % CH) * N;
      if (kk < o.Kuse) {
        const int j = S_.col[o.base + kk];
        const double x = -(xi - S_.x[j]), y = -(yi - S_.y[j]);
        const double xs = x * rhh, ys = y * rhh;
        if (role == 0) {
          mono_row<M>(xs, ys, row);
        } else {
          const double rad = sqrt(x * x + y * y) * rhh;
          abf_row<M>(xs, ys, wab(rad, S_.h2), row);
        }
      } else {
        for (int a = 0; a < N; a++) row[a] = 0.0;
      }
    }
    __syncthreads();
    if (own) {
      const int nk = min(CH, o.Kuse - c0);
      for (int k = 0; k < nk; k++) {
        const double xr = Xc[k * N + r];
#pragma unroll
        for (int m = 0; m < MC; m++) {
          const int c = q + 4 * m;
          if (c < N) acc[m] = fma(xr, Gc[k * N + c], acc[m]);
        }
      }
    }
    __syncthreads();
  }
  // order reduction near walls: rows and columns beyond nkeep become the
  // identity (labf.F90:126-148, :311-350)
  if (own) {
#pragma unroll
    for (int m = 0; m < MC; m++) {
      const int c = q + 4 * m;
      if (c < N) {
        if (r >= o.nkeep || c >= o.nkeep) acc[m] = (r == c) ? 1.0 : 0.0;
        A[r * LDA + c] = acc[m];
      }
    }
  }
  if (t < N) perm[t] = t;
  __syncthreads();

  // ---- LU with partial pivoting, rows swapped in LDS ----
  int sing = 0;
  for (int k = 0; k < N; k++) {
    if (t < 64) {
      // the largest |A(i, k)|, i >= k (lowest i on ties; NaN counts as 0)
      double best = -1.0;
      int bi = N;
      for (int i = k + t; i < N; i += 64) {
        double v = fabs(A[i * LDA + k]);
        if (v != v) v = 0.0;
        if (v > best) { best = v; bi = i; }
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const double ob = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
      }
      if (t == 0) reinterpret_cast<int *>(shd)[0] = bi;
    }
    __syncthreads();
    const int p = reinterpret_cast<int *>(shd)[0];
    if (p != k) {
      for (int c = t; c < N; c += BW) {
        const double v = A[k * LDA + c];
        A[k * LDA + c] = A[p * LDA + c];
        A[p * LDA + c] = v;
      }
      if (t == 0) {
        const int v = perm[k];
        perm[k] = perm[p];
        perm[p] = v;
      }
    }
    __syncthreads();
    const double piv = A[k * LDA + k];
    if (!(piv != 0.0)) sing = 1;
    const double rinv = 1.0 / (piv == 0.0 ? 1.0 : piv);
    for (int i = k + 1 + t; i < N; i += BW) A[i * LDA + k] *= rinv;
    __syncthreads();
    const int w = N - k - 1;
    for (int e = t; e < w * w; e += BW) {
      const int i = k + 1 + e / w, j = k + 1 + e % w;
      A[i * LDA + j] = fma(-A[i * LDA + k], A[k * LDA + j], A[i * LDA + j]);
    }
    __syncthreads();
  }

  // ---- substitutions: L U x = P v for the four right-hand sides in V ----
  // (thread (r, q): row r, right-hand side q)
  auto solve = [&](double *V) {
    for (int k = 0; k < N; k++) {
      if (own && r > k) V[q * N + r] = fma(-A[r * LDA + k], V[q * N + k], V[q * N + r]);
      __syncthreads();
    }
    for (int k = N - 1; k >= 0; k--) {
      if (own && r == k) V[q * N + k] /= A[k * LDA + k] == 0.0 ? 1.0 : A[k * LDA + k];
      __syncthreads();
      if (own && r < k) V[q * N + r] = fma(-A[r * LDA + k], V[q * N + k], V[q * N + r]);
      __syncthreads();
    }
  };
  double b4[4];
  rhs_row(r, o.sysD, o.sysF, o.mf, b4);   // row r of the right-hand sides
  if (own) {
    double bp[4];
    rhs_row(perm[r], o.sysD, o.sysF, o.mf, bp);
    Xv[q * N + r] = bp[q];
  }
  __syncthreads();
  solve(Xv);
  // residual b - A0 x of row r in double-double: the row's four threads
  // each sum their columns, then two lane exchanges combine them
  {
    double sh[4] = {0.0, 0.0, 0.0, 0.0}, sl[4] = {0.0, 0.0, 0.0, 0.0};
    if (own) {
#pragma unroll
      for (int m = 0; m < MC; m++) {
        const int c = q + 4 * m;
        if (c < N) {
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const double pr = acc[m] * Xv[v * N + c];
            const double pe = fma(acc[m], Xv[v * N + c], -pr);
            double e;
            two_sum(sh[v], pr, sh[v], e);
            sl[v] += e + pe;
          }
        }
      }
    }
#pragma unroll
    for (int off = 1; off <= 2; off <<= 1)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const double oh = __shfl_xor(sh[v], off), ol = __shfl_xor(sl[v], off);
        double e;
        two_sum(sh[v], oh, sh[v], e);
        sl[v] += e + ol;
      }
    if (own && q == 0) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        double d, de;
        two_sum(b4[v], -sh[v], d, de);
        Rv[v * N + r] = d + (de - sl[v]);
      }
    }
  }
  __syncthreads();
  // the correction: L U dx = P res, then x += dx
  double rp = 0.0;
  if (own) rp = Rv[q * N + perm[r]];
  __syncthreads();
  if (own) Rv[q * N + r] = rp;
  __syncthreads();
  solve(Rv);
  if (own) Xv[q * N + r] += Rv[q * N + r];
  if (sing && t == 0) atomicAdd(S_.singular, 1);
  __syncthreads();

  // ---- weights, one neighbour per thread; node volume ----
  double vsum = 0.0;
  for (int k = t; k < o.Kfull; k += BW) {
    const int j = S_.col[o.base + k];
    const double x = -(xi - S_.x[j]), y = -(yi - S_.y[j]);
    const double rad = sqrt(x * x + y * y) * rhh;
    const double w = wab(rad, S_.h2);
    if (o.sysD) {
      S_.wx[o.base + k] = abf_dot<M>(x * rhh, y * rhh, w, Xv) * rhh;
      S_.wy[o.base + k] = abf_dot<M>(x * rhh, y * rhh, w, Xv + N) * rhh;
      S_.wl[o.base + k] = abf_dot<M>(x * rhh, y * rhh, w, Xv + 2 * N) * rhh * rhh;
    }
    if (o.sysF) {
      S_.wh[o.base + k] = abf_dot<M>(x * rhh, y * rhh, w, Xv + 3 * N);
      vsum += (9.0 / SV_PI) * exp(-9.0 * rad * rad) / (o.hh * o.hh);
    }
  }
  if (o.sysF) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) vsum += __shfl_xor(vsum, off);
    __syncthreads();
    if ((t & 63) == 0) shd[1 + t / 64] = vsum;
    __syncthreads();
    if (t == 0) {
      double s = 0.0;
      for (int wv = 0; wv < BW / 64; wv++) s += shd[1 + wv];
      S_.vol[o.i] = 1.0 / s;
    }
  }
}

// The setup sums below run one thread per node over its CSR row, adding in
// the reference's sequential order.  A block's rows are one contiguous range
// of the CSR arrays (row_ptr is a prefix sum), which is staged through LDS in
// coalesced chunks; each thread then adds the part of its row that lies in the
// chunk.  Entries past cnt[i] (rows pruned after the lists were built) lie in
// the range and are skipped by the per-row bounds.
constexpr int SUM_B = 256;     // nodes (threads) per block
constexpr int SUM_CH = 1024;   // CSR entries per staged chunk

// calc_labf_sums (labf.F90:396-451): row sums of the four weight arrays, then
// the gradient and Laplacian weights scaled by L_char (written back from the
// staged chunk, so every CSR entry of the range is read and written once)
__global__ __launch_bounds__(SUM_B) void k_sums(int npfb, const int64_t *row_ptr, const int *cnt,
                                                double L, double *wx, double *wy, double *wl,
                                                const double *wh, double *gsx, double *gsy,
                                                double *lsum, double *hsum) {
  __shared__ double cx[SUM_CH], cy[SUM_CH], cl[SUM_CH], ch[SUM_CH];
  const int i0 = blockIdx.x * SUM_B;
  const int i1 = min(i0 + SUM_B, npfb);
  const int i = i0 + (int)threadIdx.x;
  const int64_t R0 = row_ptr[i0], R1 = row_ptr[i1];
  int64_t b = 0;
  int K = 0;
  if (i < npfb) { b = row_ptr[i]; K = cnt[i]; }
  double sh = 0.0, sx = 0.0, sy = 0.0, sl = 0.0;
  for (int64_t cs = R0; cs < R1; cs += SUM_CH) {
    const int n = (int)min((int64_t)SUM_CH, R1 - cs);
    for (int e = threadIdx.x; e < n; e += SUM_B) {
      cx[e] = wx[cs + e]; cy[e] = wy[cs + e]; cl[e] = wl[cs + e]; ch[e] = wh[cs + e];
    }
    __syncthreads();
    const int lo = (int)(max(b, cs) - cs);
    const int hi = (int)(min(b + K, cs + n) - cs);
    for (int e = lo; e < hi; e++) {
      sh += ch[e];
      sx += cx[e];
      sy += cy[e];
      sl += cl[e];
    }
    for (int e = threadIdx.x; e < n; e += SUM_B) {
      wx[cs + e] = cx[e] / L;
      wy[cs + e] = cy[e] / L;
      wl[cs + e] = cl[e] / L / L;
    }
    __syncthreads();
  }
  if (i >= npfb) return;
  hsum[i] = sh;
  gsx[i] = sx / L;
  gsy[i] = sy / L;
  lsum[i] = sl / L / L;
}

// filter_coefficients (labf.F90:1723-1811).  Two passes over the block's
// staged range: the first sums (test function) x w_hyp along each row, which
// fixes the row's coefficient; the second scales the row by it, sums the
// scaled weights and writes the chunk back.
__global__ __launch_bounds__(SUM_B) void k_filter_coeffs(int npfb, const int64_t *row_ptr,
                                                         const int *cnt, const int *col,
                                                         const double *x, const double *y,
                                                         const double *h, const double *s,
                                                         const int *nt, int nb, double *wh,
                                                         double *hsum) {
  __shared__ double cw[SUM_CH];
  __shared__ int cc[SUM_CH];
  const int i0 = blockIdx.x * SUM_B;
  const int i1 = min(i0 + SUM_B, npfb);
  const int i = i0 + (int)threadIdx.x;
  const int64_t R0 = row_ptr[i0], R1 = row_ptr[i1];
  const bool act = i < npfb;
  int64_t b = 0;
  int K = 0;
  if (act) { b = row_ptr[i]; K = cnt[i]; }
  const int t = (act && nb > 0) ? nt[i] : 999;
  // boundary nodes filter along the boundary only: radial test function,
  // length scale s (:1770-1792); the others the product test function with
  // length scale 2 h sqrt(pi / K) (:1738-1755)
  const bool radial = t >= 0 && t <= 2;
  double tmp = 0.0, xi = 0.0, yi = 0.0;
  if (act) {
    xi = x[i];
    yi = y[i];
    if (radial) {
      tmp = (2.0 / 3.0) * SV_PI / s[i];
    } else {
      const double lscal = 2.0 * h[i] * sqrt(SV_PI / (double)K);
      tmp = (2.0 / 3.0) * SV_PI / lscal;
    }
  }
  double ls = 0.0;
  for (int64_t cs = R0; cs < R1; cs += SUM_CH) {
    const int n = (int)min((int64_t)SUM_CH, R1 - cs);
    for (int e = threadIdx.x; e < n; e += SUM_B) { cw[e] = wh[cs + e]; cc[e] = col[cs + e]; }
    __syncthreads();
    const int lo = (int)(max(b, cs) - cs);
    const int hi = (int)(min(b + K, cs + n) - cs);
    for (int e = lo; e < hi; e++) {
      const int j = cc[e];
      const double dx = x[j] - xi, dy = y[j] - yi;
      const double f = radial ? 1.0 - cos(tmp * sqrt(dx * dx + dy * dy))
                              : 1.0 - cos(tmp * dx) * cos(tmp * dy);
      ls = ls + f * cw[e];
    }
    __syncthreads();
  }
  double fc = 0.0;
  if (act) {
    if (radial) {
      fc = (1.0 / 3.0) / ls;
      fc = fc * 0.5;
    } else {
      fc = (2.0 / 3.0) / ls;
      // weaker filtering on the FD rows of walls (:1756-1763)
      if (t < 0 && nt[i + t] == 0) {
        const double oosqrt2 = 1.0 / sqrt(2.0);
        if (t == -1) fc = fc * 0.5 * oosqrt2;
        if (t == -2) fc = fc * 0.5;
        if (t == -3) fc = fc * oosqrt2;
        if (t == -4) fc = fc * oosqrt2;
      }
    }
  }
  double hs = 0.0;
  for (int64_t cs = R0; cs < R1; cs += SUM_CH) {
    const int n = (int)min((int64_t)SUM_CH, R1 - cs);
    for (int e = threadIdx.x; e < n; e += SUM_B) cw[e] = wh[cs + e];
    __syncthreads();
    const int lo = (int)(max(b, cs) - cs);
    const int hi = (int)(min(b + K, cs + n) - cs);
    for (int e = lo; e < hi; e++) {
      const double w = cw[e] * fc;
      cw[e] = w;
      hs = hs + w;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n; e += SUM_B) wh[cs + e] = cw[e];
    __syncthreads();
  }
  if (act) hsum[i] = hs;
}

#define SV_TUNE_SAMPLE 4096
#define SV_TUNE_VARIANTS 4

// systems per wavefront of order m (Pack<m>::G)
int pack_of(int m) { return m == 8 ? Pack<8>::G : m == 6 ? Pack<6>::G : Pack<4>::G; }

// Units of a list sorted by order: a new unit at every change of order and
// after every G_m systems of order m.
std::vector<int> units_of(const std::vector<int> &order) {
  std::vector<int> u;
  int run = 0;
  for (size_t q = 0; q < order.size(); q++) {
    if (q == 0 || order[q] != order[q - 1] || run == pack_of(order[q])) {
      u.push_back((int)q);
      run = 0;
    }
    run++;
  }
  return u;
}

template <int M, int CH>
int launch_tiles(sv_ctx *c, const SysArgs &b, int count, bool persistent) {
  const size_t shm = sizeof(double) * Lds<M, CH>::TOTAL;
  constexpr int G = Pack<M>::G;
  const int units = (count + G - 1) / G;
  if (!persistent) {
    k_weights<M, CH><<<units, SV_WAVE, shm, c->stream>>>(b);
  } else {
    int dev = c->device, ncu = 0, occ = 0;
    SV_HIP(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    SV_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_weights_persistent<M, CH>,
                                                            SV_WAVE, shm));
    if (occ < 1) occ = 1;
    int grid = ncu * occ;
    if (grid > units) grid = units;
    SV_HIP(c, hipMemsetAsync(b.queue, 0, sizeof(int), c->stream));
    k_weights_persistent<M, CH><<<grid, SV_WAVE, shm, c->stream>>>(b);
  }
  SV_HIP(c, hipGetLastError());
  return 0;
}

// variant: bit 0 = persistent work-queue grid, bit 1 = 32-neighbour assembly
// chunks (else 16)
template <int M>
int launch_order(sv_ctx *c, SysArgs a, int first, int count, int variant) {
  if (count <= 0) return 0;
  SysArgs b = a;
  b.sys_node = a.sys_node + first;
  b.sys_kind = a.sys_kind + first;
  b.nsys = count;
  const bool persistent = (variant & 1) != 0;
  return (variant & 2) ? launch_tiles<M, 32>(c, b, count, persistent)
                       : launch_tiles<M, 16>(c, b, count, persistent);
}

template <int CH>
int launch_mixed_tiles(sv_ctx *c, const SysArgs &b, int mmax) {
  // LDS of the largest need among the orders up to mmax
  int tot = Lds<4, CH>::TOTAL;
  if (mmax >= 6) tot = std::max(tot, (int)Lds<6, CH>::TOTAL);
  if (mmax >= 8) tot = std::max(tot, (int)Lds<8, CH>::TOTAL);
  const size_t shm = sizeof(double) * tot;
  int dev = c->device, ncu = 0, occ = 0;
  SV_HIP(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  SV_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_weights_mixed<CH>, SV_WAVE,
                                                          shm));
  if (occ < 1) occ = 1;
  int grid = ncu * occ;
  if (grid > b.nunits) grid = b.nunits;
  SV_HIP(c, hipMemsetAsync(b.queue, 0, sizeof(int), c->stream));
  k_weights_mixed<CH><<<grid, SV_WAVE, shm, c->stream>>>(b);
  SV_HIP(c, hipGetLastError());
  return 0;
}

// systems of more than 64 unknowns: one block each
template <int M>
int launch_big(sv_ctx *c, SysArgs a, int first, int count) {
  if (count <= 0) return 0;
  SysArgs b = a;
  b.sys_node = a.sys_node + first;
  b.sys_kind = a.sys_kind + first;
  b.nsys = count;
  const size_t shm = sizeof(double) * Big<M>::TOTAL;
  k_weights_big<M><<<count, Big<M>::BW, shm, c->stream>>>(b);
  SV_HIP(c, hipGetLastError());
  return 0;
}

// mode 1: one mixed-order persistent launch with 16-neighbour tiles, 2: with
// 32; a.unit_first / a.nunits: the units of the count systems
int launch_mixed(sv_ctx *c, SysArgs a, int count, int mmax, int mode) {
  if (count <= 0) return 0;
  SysArgs b = a;
  b.nsys = count;
  return mode == 2 ? launch_mixed_tiles<32>(c, b, mmax) : launch_mixed_tiles<16>(c, b, mmax);
}

}  // namespace

// Build the system lists (one per order), launch, and check for zero pivots.
int sv_dev_weights(sv_ctx *c, const int *order_map_host) {
  const int n = c->npfb;
  std::vector<int> cnt(n), cs(n);
  std::vector<double> h(n), hs(n);
  const std::vector<int> &nt = c->hnt;
  SV_HIP(c, hipMemcpy(cnt.data(), c->cnt, sizeof(int) * n, hipMemcpyDeviceToHost));
  SV_HIP(c, hipMemcpy(cs.data(), c->cnt_small, sizeof(int) * n, hipMemcpyDeviceToHost));
  SV_HIP(c, hipMemcpy(h.data(), c->h, sizeof(double) * n, hipMemcpyDeviceToHost));
  SV_HIP(c, hipMemcpy(hs.data(), c->hs, sizeof(double) * n, hipMemcpyDeviceToHost));
  // systems grouped by order: entries (order, K, node, kind)
  struct Sys { int m, K, node, kind; };
  std::vector<Sys> sys;
  sys.reserve(2 * (size_t)n);
  for (int i = 0; i < n; i++) {
    const int md = order_map_host ? order_map_host[i] : c->cfg.morder;
    const int mf = order_map_host ? order_map_host[i] : c->cfg.forder;
    for (int m : {md, mf})
      if (!sv_order_supported(m))
        return sv_fail(c, SV_E_ARG, "LABFM order %d not supported on the device (4, 6, 8, 10)", m);
    // boundary nodes are not on internal_list (labf.F90:82, :268): their
    // operators come from calc_boundary_weights, which also replaces the
    // derivative weights of FD rows 1 and 2 (only their filter system is kept)
    if (c->nb > 0 && nt[i] >= 0 && nt[i] <= 2) continue;
    if (c->nb > 0 && (nt[i] == -1 || nt[i] == -2)) {
      sys.push_back({mf, cnt[i], i, SYS_F});
      continue;
    }
    if (md == mf && h[i] == hs[i] && cnt[i] == cs[i]) {
      sys.push_back({md, cnt[i], i, SYS_D | SYS_F});
    } else {
      sys.push_back({md, cs[i], i, SYS_D});
      sys.push_back({mf, cnt[i], i, SYS_F});
    }
  }
  // by order, then by decreasing stencil length, stable: a counting sort over
  // the (order, length) keys gives the order of a stable comparison sort
  {
    int kmax = 0;
    for (const Sys &e : sys) kmax = std::max(kmax, e.K);
    const int nk = kmax + 1;
    auto key = [&](const Sys &e) { return (e.m / 2 - 2) * nk + (kmax - e.K); };
    std::vector<int64_t> start((size_t)4 * nk + 1, 0);
    for (const Sys &e : sys) start[key(e) + 1]++;
    for (size_t q = 1; q < start.size(); q++) start[q] += start[q - 1];
    std::vector<Sys> sorted(sys.size());
    for (const Sys &e : sys) sorted[start[key(e)]++] = e;
    sys.swap(sorted);
  }
  std::vector<int> hn(sys.size()), hk(sys.size());
  for (size_t s = 0; s < sys.size(); s++) { hn[s] = sys[s].node; hk[s] = sys[s].kind; }
  DevTmp<int> dn, dk, dq;
  SV_HIP(c, dn.alloc(sys.size() + 1));
  SV_HIP(c, dk.alloc(sys.size() + 1));
  SV_HIP(c, dq.alloc(4));
  SV_HIP(c, hipMemcpy(dn, hn.data(), sizeof(int) * sys.size(), hipMemcpyHostToDevice));
  SV_HIP(c, hipMemcpy(dk, hk.data(), sizeof(int) * sys.size(), hipMemcpyHostToDevice));
  const size_t L = (size_t)c->nnz + 1;
  if (!c->wx) {
    if (sv_alloc(c, (void **)&c->wx, sizeof(double) * L) ||
        sv_alloc(c, (void **)&c->wy, sizeof(double) * L) ||
        sv_alloc(c, (void **)&c->wl, sizeof(double) * L) ||
        sv_alloc(c, (void **)&c->wh, sizeof(double) * L) ||
        sv_alloc(c, (void **)&c->vol, sizeof(double) * n) ||
        sv_alloc(c, (void **)&c->gsx, sizeof(double) * n) ||
        sv_alloc(c, (void **)&c->gsy, sizeof(double) * n) ||
        sv_alloc(c, (void **)&c->lsum, sizeof(double) * n) ||
        sv_alloc(c, (void **)&c->hsum, sizeof(double) * n) ||
        sv_alloc(c, (void **)&c->singular, sizeof(int) * 4))
      return SV_E_NOMEM;
  }
  SV_HIP(c, hipMemsetAsync(c->wx, 0, sizeof(double) * L, c->stream));
  SV_HIP(c, hipMemsetAsync(c->wy, 0, sizeof(double) * L, c->stream));
  SV_HIP(c, hipMemsetAsync(c->wl, 0, sizeof(double) * L, c->stream));
  SV_HIP(c, hipMemsetAsync(c->wh, 0, sizeof(double) * L, c->stream));
  SV_HIP(c, hipMemsetAsync(c->singular, 0, sizeof(int) * 4, c->stream));
  SysArgs a{dn, dk, (int)sys.size(), c->x, c->y, c->h, c->hs, c->node_type, c->row_ptr,
            c->col_raw, c->cnt, c->cnt_small, c->h2, c->wx, c->wy, c->wl, c->wh, c->vol,
            c->singular, dq, nullptr, nullptr, 0};
  EventTmp e0, e1, t0, t1;
  SV_HIP(c, e0.create());
  SV_HIP(c, e1.create());
  SV_HIP(c, t0.create());
  SV_HIP(c, t1.create());
  SV_HIP(c, hipEventRecord(e0, c->stream));
  float tune_ms = 0.0f;    // autotuning sample launches (2 passes each)
  float solve_ms = 0.0f;   // the full launches
  auto launch = [&](int m, SysArgs args, int f0, int cnt, int variant) -> int {
    if (m == 10) return launch_big<10>(c, args, f0, cnt);
    if (m == 4) return launch_order<4>(c, args, f0, cnt, variant);
    if (m == 6) return launch_order<6>(c, args, f0, cnt, variant);
    return launch_order<8>(c, args, f0, cnt, variant);
  };
  // time fn() twice (the first pass warms up), return the second
  auto timed = [&](auto fn, float &ms) -> int {
    for (int rep = 0; rep < 2; rep++) {
      SV_HIP(c, hipEventRecord(t0, c->stream));
      if (int rc = fn()) return rc;
      SV_HIP(c, hipEventRecord(t1, c->stream));
      SV_HIP(c, hipEventSynchronize(t1));
      SV_HIP(c, hipEventElapsedTime(&ms, t0, t1));
    }
    return 0;
  };
  // order slices of the (order, length)-sorted list
  struct Slice { int m, first, count; };
  std::vector<Slice> slices;
  bool has_big = false;
  for (int m : {4, 6, 8, 10}) {
    int f = 0;
    while (f < (int)sys.size() && sys[f].m != m) f++;
    int cnt = 0;
    while (f + cnt < (int)sys.size() && sys[f + cnt].m == m) cnt++;
    if (cnt) slices.push_back({m, f, cnt});
    if (cnt && m > 8) has_big = true;
  }
  const int S = SV_TUNE_SAMPLE;
  // 1. per order: the launch variant (grid kind x assembly tile), timed on a
  //    strided sample of that order's systems (sorted by stencil length, so
  //    the sample has the same length mix); the sample's weights are
  //    rewritten by the full launch with identical values
  for (const Slice &sl : slices) {
    if (sl.m > 8) continue;   // one launch shape
    const int mi = sl.m / 2 - 2;
    if (!c->weights_autotune || sl.count < 4 * S || c->wtune[mi] >= 0) continue;
    const int stride = sl.count / S;
    std::vector<int> sn(S), sk(S);
    for (int q = 0; q < S; q++) {
      sn[q] = hn[sl.first + q * stride];
      sk[q] = hk[sl.first + q * stride];
    }
    DevTmp<int> tn, tk;
    SV_HIP(c, tn.alloc(S));
    SV_HIP(c, tk.alloc(S));
    SV_HIP(c, hipMemcpy(tn, sn.data(), sizeof(int) * S, hipMemcpyHostToDevice));
    SV_HIP(c, hipMemcpy(tk, sk.data(), sizeof(int) * S, hipMemcpyHostToDevice));
    SysArgs ts = a;
    ts.sys_node = tn;
    ts.sys_kind = tk;
    float best = 1e30f;
    for (int v = 0; v < SV_TUNE_VARIANTS; v++) {
      float tv = 0.0f;
      if (int rc = timed([&] { return launch(sl.m, ts, 0, S, v); }, tv)) return rc;
      c->wtune_ms[mi][v] = tv;
      tune_ms += 2.0f * tv;
      if (tv < best) { best = tv; c->wtune[mi] = v; }
    }
  }
  auto order_variant = [&](int m) {
    if (m > 8) return 0;
    const int mi = m / 2 - 2;
    return (c->weights_autotune && c->wtune[mi] >= 0) ? c->wtune[mi] : c->weights_fixed;
  };
  // 2. several orders: one persistent launch over all of them (queue order 8
  //    first) against the per-order launches, on a strided sample of that queue
  int mode = c->weights_autotune ? 0 : c->weights_mixed;
  std::vector<int> qn, qk, qm;   // the mixed queue
  if (has_big) {
    mode = 0;   // the mixed launch covers orders 4-8
  } else if (slices.size() >= 2) {
    for (auto it = slices.rbegin(); it != slices.rend(); ++it)
      for (int q = it->first; q < it->first + it->count; q++) {
        qn.push_back(hn[q]);
        qk.push_back(hk[q]);
        qm.push_back(it->m);
      }
    const int total = (int)qn.size();
    if (c->weights_autotune && total >= 4 * S && c->wmixed < 0) {
      const int stride = total / S;
      std::vector<int> sn(S), sk(S), sm(S);
      for (int q = 0; q < S; q++) {
        sn[q] = qn[q * stride];
        sk[q] = qk[q * stride];
        sm[q] = qm[q * stride];
      }
      const std::vector<int> su = units_of(sm);
      DevTmp<int> tn, tk, tm, tu;
      SV_HIP(c, tn.alloc(S));
      SV_HIP(c, tk.alloc(S));
      SV_HIP(c, tm.alloc(S));
      SV_HIP(c, tu.alloc(su.size()));
      SV_HIP(c, hipMemcpy(tn, sn.data(), sizeof(int) * S, hipMemcpyHostToDevice));
      SV_HIP(c, hipMemcpy(tk, sk.data(), sizeof(int) * S, hipMemcpyHostToDevice));
      SV_HIP(c, hipMemcpy(tm, sm.data(), sizeof(int) * S, hipMemcpyHostToDevice));
      SV_HIP(c, hipMemcpy(tu, su.data(), sizeof(int) * su.size(), hipMemcpyHostToDevice));
      SysArgs ts = a;
      ts.sys_node = tn;
      ts.sys_kind = tk;
      ts.sys_order = tm;
      ts.unit_first = tu;
      ts.nunits = (int)su.size();
      float tv[3] = {0.0f, 0.0f, 0.0f};
      // per-order launches over the sample's contiguous order groups
      if (int rc = timed([&] {
            int f = 0;
            while (f < S) {
              int n2 = 0;
              while (f + n2 < S && sm[f + n2] == sm[f]) n2++;
              if (int r2 = launch(sm[f], ts, f, n2, order_variant(sm[f]))) return r2;
              f += n2;
            }
            return 0;
          }, tv[0]))
        return rc;
      const int mmax = qm[0];
      for (int md = 1; md <= 2; md++)
        if (int rc = timed([&] { return launch_mixed(c, ts, S, mmax, md); }, tv[md])) return rc;
      int bestm = 0;
      for (int md = 0; md < 3; md++) {
        c->wmixed_ms[md] = tv[md];
        tune_ms += 2.0f * tv[md];
        if (tv[md] < tv[bestm]) bestm = md;
      }
      c->wmixed = bestm;
    }
    if (c->weights_autotune && c->wmixed >= 0) mode = c->wmixed;
  } else {
    mode = 0;
  }
  // 3. the launches
  c->wmixed_used = mode;
  for (int k = 0; k < 3; k++) c->wvariant[k] = -1;
  if (mode == 0) {
    for (const Slice &sl : slices) {
      const int variant = order_variant(sl.m);
      float tl = 0.0f;
      SV_HIP(c, hipEventRecord(t0, c->stream));
      if (int rc = launch(sl.m, a, sl.first, sl.count, variant)) return rc;
      SV_HIP(c, hipEventRecord(t1, c->stream));
      SV_HIP(c, hipEventSynchronize(t1));
      SV_HIP(c, hipEventElapsedTime(&tl, t0, t1));
      solve_ms += tl;
      if (sl.m <= 8) c->wvariant[sl.m / 2 - 2] = variant;
    }
  } else {
    const int total = (int)qn.size();
    const std::vector<int> qu = units_of(qm);
    DevTmp<int> mn, mk, mm, mu;
    SV_HIP(c, mn.alloc(total));
    SV_HIP(c, mk.alloc(total));
    SV_HIP(c, mm.alloc(total));
    SV_HIP(c, mu.alloc(qu.size()));
    SV_HIP(c, hipMemcpy(mn, qn.data(), sizeof(int) * total, hipMemcpyHostToDevice));
    SV_HIP(c, hipMemcpy(mk, qk.data(), sizeof(int) * total, hipMemcpyHostToDevice));
    SV_HIP(c, hipMemcpy(mm, qm.data(), sizeof(int) * total, hipMemcpyHostToDevice));
    SV_HIP(c, hipMemcpy(mu, qu.data(), sizeof(int) * qu.size(), hipMemcpyHostToDevice));
    SysArgs am = a;
    am.sys_node = mn;
    am.sys_kind = mk;
    am.sys_order = mm;
    am.unit_first = mu;
    am.nunits = (int)qu.size();
    float tl = 0.0f;
    SV_HIP(c, hipEventRecord(t0, c->stream));
    if (int rc = launch_mixed(c, am, total, qm[0], mode)) return rc;
    SV_HIP(c, hipEventRecord(t1, c->stream));
    SV_HIP(c, hipEventSynchronize(t1));
    SV_HIP(c, hipEventElapsedTime(&tl, t0, t1));
    solve_ms += tl;
  }
  SV_HIP(c, hipEventRecord(e1, c->stream));
  SV_HIP(c, hipStreamSynchronize(c->stream));
  float ms = 0;
  SV_HIP(c, hipEventElapsedTime(&ms, e0, e1));
  (void)ms;
  c->t_weights_tune = tune_ms;
  c->t_weights = solve_ms;
  c->n_systems = (int)sys.size();
  int nsing = 0;
  SV_HIP(c, hipMemcpy(&nsing, c->singular, sizeof(int), hipMemcpyDeviceToHost));
  c->have_weights = true;
  c->have_sums = false;
  c->have_filter = false;
  c->sell_dirty = true;
  c->L.valid = false;
  c->n_singular = nsing;
  return 0;
}

int sv_dev_sums(sv_ctx *c) {
  const int n = c->npfb, B = SUM_B;
  k_sums<<<(n + B - 1) / B, B, 0, c->stream>>>(n, c->row_ptr, c->cnt, c->L_char, c->wx, c->wy,
                                               c->wl, c->wh, c->gsx, c->gsy, c->lsum, c->hsum);
  SV_HIP(c, hipGetLastError());
  if (int rc = sv_dev_boundary_tt_sums(c)) return rc;
  SV_HIP(c, hipStreamSynchronize(c->stream));
  c->have_sums = true;
  c->sell_dirty = true;
  c->L.valid = false;
  return 0;
}

int sv_dev_filter_coeffs(sv_ctx *c) {
  const int n = c->npfb, B = SUM_B;
  k_filter_coeffs<<<(n + B - 1) / B, B, 0, c->stream>>>(n, c->row_ptr, c->cnt, c->col_raw, c->x,
                                                        c->y, c->h, c->s, c->node_type, c->nb,
                                                        c->wh, c->hsum);
  SV_HIP(c, hipGetLastError());
  SV_HIP(c, hipStreamSynchronize(c->stream));
  c->have_filter = true;
  c->sell_dirty = true;
  c->L.valid = false;
  return 0;
}

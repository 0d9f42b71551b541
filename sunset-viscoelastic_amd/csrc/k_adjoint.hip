// k_adjoint.hip -- hand-written reverse mode of the direct-form (DI) RK3 step
// at a fixed time step: the vector-Jacobian product of calc_all_rhs
// (rhs_di.F90:47-476 with the full-gradient reading of calc_rhs_c), of the
// four RK stages and registers (step.F90:281-409), of the hyperviscous filter
// (calc_filtered_var, derivatives.F90:360-414) and of the mass correction
// (rhs_di.F90:667-720), and checkpointed rollouts of a trace loss.
//
// Stencil transposes without atomics: every forward link (i -> j, weights w)
// is stored once more in a transposed SELL-64 layout keyed by j, so the
// adjoint of "D_i = sum_k w_ik G_j - G_i sum_k w_ik" is a gather
//   Gbar_j += sum_{(i,k): j_k = i's link} w_ik Dbar_i
// over the transposed list of j in a fixed order (ascending i): deterministic.
//
// Each RK stage's adjoint has two parts:
//   pointwise    Rbar_i and the stage tape (own gather record G_i and
//                derivative values D_i written by the forward stage) -> the
//                adjoints of the 14 derivative values Dbar_i (+3 Laplacians of
//                c with Mdiff) and the own-node adjoint of G_i
//   transposed   gather of Dbar, G -> U (u = rou/ro), and the RK register
//                recurrences that give the previous stage's Rbar
// The pointwise part of stage k-1 needs only node i's own Rbar, so it runs
// inside the transposed gather of stage k (two buffer sets, ping-pong), and
// the pointwise part of stage 4 inside the fused mass-correction + filter
// adjoint: one step's adjoint is six launches (a per-segment sum, the filter,
// four transposed gathers).  Rollouts keep the state every `checkpoint`
// steps; the adjoint recomputes one checkpoint interval with the tape on and
// walks it backwards.
#include <string.h>

#include "sv_internal.h"

namespace {

constexpr int BS = SV_BLOCK;

#ifndef SV_ADJ_BATCH
#define SV_ADJ_BATCH 4     // transposed links per batch of requests
#endif
// The reverse sweeps run blocks of AB threads, BS/AB of them per 256-node
// block of the node layout, so that small node sets (a 128x128 lattice is 64
// such blocks) still occupy every CU.
#ifndef SV_ADJ_THREADS
#define SV_ADJ_THREADS 64
#endif
constexpr int AB = SV_ADJ_THREADS;
constexpr int SUBS = BS / AB;
static_assert(BS % AB == 0 && AB % 64 == 0, "adjoint blocks are whole waves dividing SV_BLOCK");

__device__ __forceinline__ void ld6(const double *p, double f[6]) {
  const double2 *q = reinterpret_cast<const double2 *>(p);
  const double2 a = q[0], b = q[1], c = q[2];
  f[0] = a.x; f[1] = a.y; f[2] = b.x; f[3] = b.y; f[4] = c.x; f[5] = c.y;
}

__device__ __forceinline__ void st6(double *p, const double f[6]) {
  double2 *q = reinterpret_cast<double2 *>(p);
  q[0] = double2{f[0], f[1]};
  q[1] = double2{f[2], f[3]};
  q[2] = double2{f[4], f[5]};
}

struct NodeRef {
  int sg, i, ii;
  bool act;
};

// node of this thread in sub-block sb (AB threads) of the node layout
__device__ __forceinline__ NodeRef node_ref_sub(const int *blk_seg, const sv_segk *seg, int sb) {
  NodeRef r;
  const int bid = sb / SUBS;
  r.sg = blk_seg[bid];
  const sv_segk &S = seg[r.sg];
  const int li = (bid - S.blk0) * BS + (sb % SUBS) * AB + (int)threadIdx.x;
  r.act = li < S.npfb;
  r.i = S.base + li;
  r.ii = r.act ? r.i : S.base;
  return r;
}

// node of this thread when LPN lanes share a node (LPN = 1 or 4): a block of
// AB threads covers AB / LPN consecutive nodes, BS * LPN / AB blocks per
// 256-node block of the node layout
template <int LPN>
__device__ __forceinline__ NodeRef node_ref_lanes(const int *blk_seg, const sv_segk *seg,
                                                  int sb) {
  if (LPN == 1) return node_ref_sub(blk_seg, seg, sb);
  constexpr int NPB = AB / LPN, SUBL = BS / NPB;
  NodeRef r;
  const int bid = sb / SUBL;
  r.sg = blk_seg[bid];
  const sv_segk &S = seg[r.sg];
  const int li = (bid - S.blk0) * BS + (sb % SUBL) * NPB + (int)threadIdx.x / LPN;
  r.act = li < S.npfb;
  r.i = S.base + li;
  r.ii = r.act ? r.i : S.base;
  return r;
}

// sum of v over the 4 lanes of a quad (DPP quad_perm [1,0,3,2] then
// [2,3,0,1]); every lane of the quad ends with the same value
__device__ __forceinline__ double quad_sum(double v) {
  int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0xB1, 0xF, 0xF, false);
  int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0xB1, 0xF, 0xF, false);
  v = v + __hiloint2double(hi, lo);
  lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x4E, 0xF, 0xF, false);
  hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x4E, 0xF, 0xF, false);
  return v + __hiloint2double(hi, lo);
}

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

// ---- pointwise adjoint of the DI right-hand side ------------------------
// Dbar record [ntot][16]: rx ry | ux uy lu | vx vy lv | cxx_x cxx_y | cxy_x
// cxy_y | cyy_x cyy_y | 0 0; with Mdiff a second record [ntot][4]: the
// Laplacian adjoints of cxx, cxy, cyy.
struct PointArgs {
  const int *blk_seg;
  const sv_segk *seg;
  const double *tG, *tD;     // stage tape
  const double *Rbar;        // [ntot][SV_US]
  const double *gsx, *gsy, *lsum;
  double *dbar, *dbarm, *own;
  double *pacc;              // [ntot][4] parameter-gradient accumulators (or null):
                             // d/d(visc_solvent), d/d(cp), d/d(lambda), d/d(eps)
};

// the pointwise adjoint of one (active) node, Rbar given in registers
template <bool MDIFF>
__device__ __forceinline__ void adj_point_node(const PointArgs &a, const NodeRef &nr,
                                               const double rb[6]) {
  const int i = nr.i;
  const sv_physk &P = a.seg[nr.sg].pk;
  double g[6];
  ld6(a.tG + (size_t)i * SV_US, g);
  const double ro = g[0], u = g[1], v = g[2], cxx = g[3], cxy = g[4], cyy = g[5];
  const double *D = a.tD + (size_t)i * SV_TAPE_D;
  const double rx = D[0], ry = D[1], ux = D[2], uy = D[3], vx = D[5], vy = D[6];
  const double gxx_x = D[8], gxx_y = D[9], gxy_x = D[10], gxy_y = D[11];
  const double gyy_x = D[12], gyy_y = D[13];
  const double Ab = rb[0], Bu = rb[1], Bv = rb[2], Xb = rb[3], Yb = rb[4], Zb = rb[5];
  // forward intermediates the adjoint needs
  const double r_ro = -ro * (ux + vy) - (u * rx + v * ry);
  const double fr = -(1.0 - P.eps * 2.0 + P.eps * (cxx + cyy + 1.0)) / P.lambda;
  double bro = 0, bu = 0, bv = 0, bcxx = 0, bcxy = 0, bcyy = 0;
  double brx = 0, bry = 0, bux = 0, buy = 0, blu = 0, bvx = 0, bvy = 0, blv = 0;
  double bgxx_x = 0, bgxx_y = 0, bgxy_x = 0, bgxy_y = 0, bgyy_x = 0, bgyy_y = 0;
  // r_rou = -tsu - csq*rx + ro*bfx + fx + cs*lu + cp*(cxx_x + cxy_y)
  const double btsu = -Bu, btsv = -Bv;
  brx += -P.csq * Bu;
  bry += -P.csq * Bv;
  bro += P.bfx * Bu + P.bfy * Bv;
  blu += P.coef_solvent * Bu;
  blv += P.coef_solvent * Bv;
  bgxx_x += P.coef_polymeric * Bu;
  bgxy_y += P.coef_polymeric * Bu;
  bgxy_x += P.coef_polymeric * Bv;
  bgyy_y += P.coef_polymeric * Bv;
  // tsu = ro*(u*ux + v*uy) - u*r_ro ; tsv = ro*(u*vx + v*vy) - v*r_ro
  bro += (u * ux + v * uy) * btsu + (u * vx + v * vy) * btsv;
  bu += ro * ux * btsu - r_ro * btsu + ro * vx * btsv;
  bv += ro * uy * btsu + ro * vy * btsv - r_ro * btsv;
  bux += ro * u * btsu;
  buy += ro * v * btsu;
  bvx += ro * u * btsv;
  bvy += ro * v * btsv;
  const double Ar = Ab - u * btsu - v * btsv;
  // r_ro = -ro*(ux + vy) - (u*rx + v*ry)
  bro += -(ux + vy) * Ar;
  bux += -ro * Ar;
  bvy += -ro * Ar;
  bu += -rx * Ar;
  brx += -u * Ar;
  bv += -ry * Ar;
  bry += -v * Ar;
  // advection of c
  bu += -(gxx_x * Xb + gxy_x * Yb + gyy_x * Zb);
  bv += -(gxx_y * Xb + gxy_y * Yb + gyy_y * Zb);
  bgxx_x += -u * Xb; bgxx_y += -v * Xb;
  bgxy_x += -u * Yb; bgxy_y += -v * Yb;
  bgyy_x += -u * Zb; bgyy_y += -v * Zb;
  // upper-convected terms
  bux += 2.0 * cxx * Xb + cxy * Yb;
  buy += 2.0 * cxy * Xb + cyy * Yb;
  bvx += cxx * Yb + 2.0 * cxy * Zb;
  bvy += cxy * Yb + 2.0 * cyy * Zb;
  bcxx += 2.0 * ux * Xb + vx * Yb;
  bcxy += 2.0 * uy * Xb + (ux + vy) * Yb + 2.0 * vx * Zb;
  bcyy += uy * Yb + 2.0 * vy * Zb;
  // relaxation fr*(c - I), fr = -(1 - 2 eps + eps*(cxx + cyy + 1))/lambda
  bcxx += fr * Xb;
  bcxy += fr * Yb;
  bcyy += fr * Zb;
  const double bfr = (cxx - 1.0) * Xb + cxy * Yb + (cyy - 1.0) * Zb;
  const double dfr = -P.eps / P.lambda * bfr;
  bcxx += dfr;
  bcyy += dfr;
  double blxx = 0, blxy = 0, blyy = 0;
  if (MDIFF) { blxx = P.Mdiff * Xb; blxy = P.Mdiff * Yb; blyy = P.Mdiff * Zb; }
  // D_i = sum_k w G_j - G_i * sum: own-node share
  const double sx = a.gsx[i], sy = a.gsy[i], sl = a.lsum[i];
  double o[6];
  o[0] = bro - (sx * brx + sy * bry);
  o[1] = bu - (sx * bux + sy * buy + sl * blu);
  o[2] = bv - (sx * bvx + sy * bvy + sl * blv);
  o[3] = bcxx - (sx * bgxx_x + sy * bgxx_y + sl * blxx);
  o[4] = bcxy - (sx * bgxy_x + sy * bgxy_y + sl * blxy);
  o[5] = bcyy - (sx * bgyy_x + sy * bgyy_y + sl * blyy);
  st6(a.own + (size_t)i * SV_US, o);
  double2 *T = reinterpret_cast<double2 *>(a.dbar + (size_t)i * SV_TAPE_D);
  T[0] = double2{brx, bry}; T[1] = double2{bux, buy}; T[2] = double2{blu, bvx};
  T[3] = double2{bvy, blv}; T[4] = double2{bgxx_x, bgxx_y}; T[5] = double2{bgxy_x, bgxy_y};
  T[6] = double2{bgyy_x, bgyy_y}; T[7] = double2{0.0, 0.0};
  if (MDIFF) {
    double2 *M = reinterpret_cast<double2 *>(a.dbarm + (size_t)i * 4);
    M[0] = double2{blxx, blxy};
    M[1] = double2{blyy, 0.0};
  }
  if (a.pacc) {
    // r_rou, r_rov hold visc_solvent*(lu, lv) and cp*(div c); the conformation
    // source fr*(c - I) holds lambda and eps through fr
    const double lu = D[4], lv = D[7];
    double2 *Pq = reinterpret_cast<double2 *>(a.pacc + (size_t)i * 4);
    double2 p0 = Pq[0], p1 = Pq[1];
    p0.x += Bu * lu + Bv * lv;
    p0.y += Bu * (gxx_x + gxy_y) + Bv * (gxy_x + gyy_y);
    p1.x += bfr * (-fr / P.lambda);
    p1.y += bfr * (-(cxx + cyy - 1.0) / P.lambda);
    Pq[0] = p0;
    Pq[1] = p1;
  }
}

template <bool MDIFF>
__global__ __launch_bounds__(AB) void k_adj_point(PointArgs a) {
  const NodeRef nr = node_ref_sub(a.blk_seg, a.seg, sv_xcd_block());
  if (!nr.act) return;
  double rb[6];
  ld6(a.Rbar + (size_t)nr.i * SV_US, rb);
  adj_point_node<MDIFF>(a, nr, rb);
}

// ---- transposed gather + G->U + RK register recurrences ------------------
struct TransArgs {
  const int *blk_seg;
  const sv_segk *seg;
  const int64_t *toff;       // transposed SELL
  const int *tcnt, *tcol;
  const double *twx, *twy, *twl;
  const double *dbar, *dbarm, *own;
  const double *tG;          // stage tape (ro, u, v of node j)
  const sv_ctrl *ctrl;       // RK coefficients (dt included)
  double *Ubar;              // STAGE 1 / 0: output
  double *Rbar;              // STAGE 2..4: next (earlier) stage's Rbar
  double *regbar;            // STAGE 2..4: in/out
  double seed;               // STAGE 1: loss seed added to cxx and cyy
  double *psum;              // STAGE 1: per-sub-block sum of Ubar_ro (mass correction)
  PointArgs nxt;             // STAGE 2..4: the pointwise adjoint of stage STAGE-1,
                             // fused (its Rbar stays in registers)
};

// one transposed link: the stencil-adjoint record d (+ m with Mdiff) of node i
// weighted by the link's gradient and Laplacian weights
template <bool MDIFF>
__device__ __forceinline__ void trans_link(double acc[6], double wx, double wy, double wl,
                                           const double2 d[7], const double2 m[2]) {
  acc[0] += wx * d[0].x + wy * d[0].y;
  acc[1] += wx * d[1].x + wy * d[1].y + wl * d[2].x;
  acc[2] += wx * d[2].y + wy * d[3].x + wl * d[3].y;
  acc[3] += wx * d[4].x + wy * d[4].y;
  acc[4] += wx * d[5].x + wy * d[5].y;
  acc[5] += wx * d[6].x + wy * d[6].y;
  if (MDIFF) {
    acc[3] += wl * m[0].x;
    acc[4] += wl * m[0].y;
    acc[5] += wl * m[1].x;
  }
}

// STAGE 0: plain vector-Jacobian product of calc_all_rhs (Ubar = J^T Rbar).
// LPN lanes per node: lane r of a node sums its links r, r + LPN, ... and the
// quad's partial sums are added in a fixed order (small node sets, where one
// lane per node leaves most of the chip idle and each lane walks ~70 links).
template <int STAGE, bool MDIFF, int LPN>
__global__ __launch_bounds__(AB) void k_adj_trans(TransArgs a) {
  __shared__ double red[AB];
  const int bid = sv_xcd_block();
  const NodeRef nr = node_ref_lanes<LPN>(a.blk_seg, a.seg, bid);
  const int j = nr.i, jj = nr.ii;
  const int lr = (int)threadIdx.x % LPN;
  const int64_t b = a.toff[jj >> 6] + (jj & 63);
  const int K = nr.act ? a.tcnt[jj] : 0;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  // links in batches of SV_ADJ_BATCH: every index and weight of a batch, then
  // every gathered record, are requested before the first sum, so the
  // dependent index -> record loads cost one memory round trip per batch
  // rather than per link (the sums keep the link order)
  int k = lr;
  for (; k + (SV_ADJ_BATCH - 1) * LPN < K; k += SV_ADJ_BATCH * LPN) {
    int ix[SV_ADJ_BATCH];
    double wx[SV_ADJ_BATCH], wy[SV_ADJ_BATCH], wl[SV_ADJ_BATCH];
#pragma unroll
    for (int u = 0; u < SV_ADJ_BATCH; u++) {
      const int64_t e = b + (int64_t)(k + u * LPN) * SV_WAVE;
      ix[u] = a.tcol[e];
      wx[u] = a.twx[e];
      wy[u] = a.twy[e];
      wl[u] = a.twl[e];
    }
    double2 d[SV_ADJ_BATCH][7], m[SV_ADJ_BATCH][2];
#pragma unroll
    for (int u = 0; u < SV_ADJ_BATCH; u++) {
      const double2 *r = reinterpret_cast<const double2 *>(a.dbar + (size_t)ix[u] * SV_TAPE_D);
#pragma unroll
      for (int q = 0; q < 7; q++) d[u][q] = r[q];
      if (MDIFF) {
        const double2 *rm = reinterpret_cast<const double2 *>(a.dbarm + (size_t)ix[u] * 4);
        m[u][0] = rm[0];
        m[u][1] = rm[1];
      }
    }
#pragma unroll
    for (int u = 0; u < SV_ADJ_BATCH; u++)
      trans_link<MDIFF>(acc, wx[u], wy[u], wl[u], d[u], m[u]);
  }
  for (; k < K; k += LPN) {
    const int64_t e = b + (int64_t)k * SV_WAVE;
    const int i = a.tcol[e];
    const double2 *r = reinterpret_cast<const double2 *>(a.dbar + (size_t)i * SV_TAPE_D);
    double2 d[7], m[2];
#pragma unroll
    for (int q = 0; q < 7; q++) d[q] = r[q];
    if (MDIFF) {
      const double2 *rm = reinterpret_cast<const double2 *>(a.dbarm + (size_t)i * 4);
      m[0] = rm[0];
      m[1] = rm[1];
    }
    trans_link<MDIFF>(acc, a.twx[e], a.twy[e], a.twl[e], d, m);
  }
  if (LPN == 4) {
#pragma unroll
    for (int f = 0; f < 6; f++) acc[f] = quad_sum(acc[f]);
  }
  double part = 0.0;
  if (nr.act && lr == 0) {
    double o[6], g[6];
    ld6(a.own + (size_t)j * SV_US, o);
    ld6(a.tG + (size_t)j * SV_US, g);
    double G[6];
#pragma unroll
    for (int f = 0; f < 6; f++) G[f] = o[f] + acc[f];
    // G = (ro, rou/ro, rov/ro, c) -> U
    const double ooro = 1.0 / g[0];
    double Ub[6];
    Ub[0] = G[0] - (G[1] * g[1] + G[2] * g[2]) * ooro;
    Ub[1] = G[1] * ooro;
    Ub[2] = G[2] * ooro;
    Ub[3] = G[3]; Ub[4] = G[4]; Ub[5] = G[5];
    const sv_ctrl *C = a.ctrl + nr.sg;
    if (STAGE == 0) {
      st6(a.Ubar + (size_t)j * SV_US, Ub);
    } else if (STAGE == 1) {
      double rg[6];
      ld6(a.regbar + (size_t)j * SV_US, rg);
#pragma unroll
      for (int f = 0; f < 6; f++) Ub[f] += rg[f];
      Ub[3] += a.seed;
      Ub[5] += a.seed;
      st6(a.Ubar + (size_t)j * SV_US, Ub);
      part = Ub[0];
    } else {
      // U_k = reg_{k-2} + a_{k-1} R_{k-1};  reg_{k-1} = reg_{k-2} + b_{k-1} R_{k-1}
      const double ra = C->RKa[STAGE - 2], rbk = C->RKb[STAGE - 2];
      double rg[6], Rn[6];
      ld6(a.regbar + (size_t)j * SV_US, rg);
#pragma unroll
      for (int f = 0; f < 6; f++) {
        Rn[f] = ra * Ub[f] + rbk * rg[f];
        rg[f] = Ub[f] + rg[f];
      }
      st6(a.regbar + (size_t)j * SV_US, rg);
      adj_point_node<MDIFF>(a.nxt, nr, Rn);
    }
  }
  if (STAGE == 1) {
    red[threadIdx.x] = part;
    __syncthreads();
    for (int s = AB / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) a.psum[bid] = red[0];
  }
}

// per-block sums of Ubar_ro (the start of a backward sweep), in the block
// layout of the LPN sweeps that write psum afterwards
template <int LPN>
__global__ __launch_bounds__(AB) void k_adj_sum(const int *blk_seg, const sv_segk *seg,
                                                const double *Ubar, double *psum) {
  __shared__ double red[AB];
  const int bid = sv_xcd_block();
  const NodeRef nr = node_ref_lanes<LPN>(blk_seg, seg, bid);
  red[threadIdx.x] = (nr.act && (int)threadIdx.x % LPN == 0) ? Ubar[(size_t)nr.i * SV_US] : 0.0;
  __syncthreads();
  for (int s = AB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) psum[bid] = red[0];
}

// mass-correction adjoint: sbar[seg] = sum_j Ubar_ro,j / sum_j vol_j (one block
// per segment, over the segment's sub-block sums), then Ubar_ro,i -= vol_i * sbar
__global__ __launch_bounds__(1024) void k_adj_sum_reduce(const sv_segk *seg, const double *psum,
                                                         int subl, double *sbar) {
  __shared__ double s1[1024];
  const sv_segk &S = seg[blockIdx.x];
  double a = 0.0;
  for (int k = threadIdx.x; k < S.nblk * subl; k += 1024) a += psum[S.blk0 * subl + k];
  s1[threadIdx.x] = a;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) s1[threadIdx.x] += s1[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) sbar[blockIdx.x] = s1[0] / S.volsum;
}

// mass-correction + filter adjoint, fused with the pointwise adjoint of
// stage 4: Uf_i = Ubar_i - [ro] vol_i sbar (evaluated for every gathered i),
// Ustar_j = Uf_j - hsum_j Uf_j + sum_t wh_t Uf_{i_t}; regbar_3 = Ustar,
// Rbar_4 = b_4 Ustar (registers) -> stage-4 pointwise adjoint
struct FiltArgs {
  const int *blk_seg;
  const sv_segk *seg;
  const int64_t *toff;
  const int *tcnt, *tcol;
  const double *twh, *hsum, *vol, *sbar;
  const double *Ubar;
  const sv_ctrl *ctrl;
  double *regbar;
  PointArgs p4;
};

template <bool MDIFF, int LPN>
__global__ __launch_bounds__(AB) void k_adj_filter(FiltArgs a) {
  const NodeRef nr = node_ref_lanes<LPN>(a.blk_seg, a.seg, sv_xcd_block());
  const int j = nr.i, jj = nr.ii;
  const int lr = (int)threadIdx.x % LPN;
  const int64_t b = a.toff[jj >> 6] + (jj & 63);
  const int K = nr.act ? a.tcnt[jj] : 0;
  const double sv = a.sbar[nr.sg];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  int k = lr;
  // batched as in k_adj_trans, lane lr of a node taking links lr, lr + LPN, ...
  for (; k + (SV_ADJ_BATCH - 1) * LPN < K; k += SV_ADJ_BATCH * LPN) {
    int ix[SV_ADJ_BATCH];
    double w[SV_ADJ_BATCH], f[SV_ADJ_BATCH][6], vo[SV_ADJ_BATCH];
#pragma unroll
    for (int u = 0; u < SV_ADJ_BATCH; u++) {
      const int64_t e = b + (int64_t)(k + u * LPN) * SV_WAVE;
      ix[u] = a.tcol[e];
      w[u] = a.twh[e];
    }
#pragma unroll
    for (int u = 0; u < SV_ADJ_BATCH; u++) {
      ld6(a.Ubar + (size_t)ix[u] * SV_US, f[u]);
      vo[u] = a.vol[ix[u]];
    }
#pragma unroll
    for (int u = 0; u < SV_ADJ_BATCH; u++) {
      f[u][0] -= vo[u] * sv;
#pragma unroll
      for (int q = 0; q < 6; q++) acc[q] += w[u] * f[u][q];
    }
  }
  for (; k < K; k += LPN) {
    const int64_t e = b + (int64_t)k * SV_WAVE;
    const int i = a.tcol[e];
    const double w = a.twh[e];
    double f[6];
    ld6(a.Ubar + (size_t)i * SV_US, f);
    f[0] -= a.vol[i] * sv;
#pragma unroll
    for (int q = 0; q < 6; q++) acc[q] += w * f[q];
  }
  if (LPN == 4) {
#pragma unroll
    for (int q = 0; q < 6; q++) acc[q] = quad_sum(acc[q]);
  }
  if (!nr.act || lr != 0) return;
  double f[6], out[6], rb[6];
  ld6(a.Ubar + (size_t)j * SV_US, f);
  f[0] -= a.vol[j] * sv;
  const double hs = a.hsum[j];
  const double b4 = a.ctrl[nr.sg].RKb[3];
#pragma unroll
  for (int q = 0; q < 6; q++) {
    out[q] = f[q] - hs * f[q] + acc[q];
    rb[q] = b4 * out[q];
  }
  st6(a.regbar + (size_t)j * SV_US, out);
  adj_point_node<MDIFF>(a.p4, nr, rb);
}

// trace loss of the state: per-block sums of (cxx + cyy - 2), then one block
// per segment adds mean/T into lacc[seg]
__global__ __launch_bounds__(BS) void k_loss_part(const int *blk_seg, const sv_segk *seg,
                                                  const double *U, double *part) {
  __shared__ double red[BS];
  const int bid = sv_xcd_block();
  const NodeRef nr = node_ref(blk_seg, seg, bid);
  double v = 0.0;
  if (nr.act) {
    const double *r = U + (size_t)nr.i * SV_US;
    v = r[3] + r[5] - 2.0;
  }
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[bid] = red[0];
}

__global__ __launch_bounds__(1024) void k_loss_acc(const sv_segk *seg, const double *part,
                                                   double *lacc) {
  __shared__ double s1[1024];
  const sv_segk &S = seg[blockIdx.x];
  double a = 0.0;
  for (int k = threadIdx.x; k < S.nblk; k += 1024) a += part[S.blk0 + k];
  s1[threadIdx.x] = a;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) s1[threadIdx.x] += s1[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) lacc[blockIdx.x] += s1[0] / S.npfb;
}

// sums of the parameter-gradient accumulators over the real nodes (one block,
// fixed order: deterministic)
__global__ __launch_bounds__(1024) void k_param_reduce(const double *pacc, int n, double *out) {
  __shared__ double s1[4][1024];
  double a[4] = {0, 0, 0, 0};
  for (int k = threadIdx.x; k < n; k += 1024)
    for (int q = 0; q < 4; q++) a[q] += pacc[(size_t)k * 4 + q];
  for (int q = 0; q < 4; q++) s1[q][threadIdx.x] = a[q];
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int q = 0; q < 4; q++) s1[q][threadIdx.x] += s1[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 4) out[threadIdx.x] = s1[threadIdx.x][0];
}

__global__ void k_zero(double *p, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0.0;
}

}  // namespace

// ---- host side ------------------------------------------------------------
struct sv_adj {
  int every = 20;            // checkpoint interval (steps)
  // transposed stencil (SELL-64 over j)
  int64_t *toff = nullptr;
  int *tcnt = nullptr, *tcol = nullptr;
  double *twx = nullptr, *twy = nullptr, *twl = nullptr, *twh = nullptr;
  int64_t tnnz = 0;
  // adjoint records [ntot][SV_US] and stencil-adjoint records
  double *Ubar = nullptr, *Rbar = nullptr, *regbar = nullptr;
  // stencil-adjoint records of two consecutive stages (ping-pong: a fused
  // kernel reads stage k's while writing stage k-1's)
  double *own[2] = {nullptr, nullptr}, *dbar[2] = {nullptr, nullptr};
  double *dbarm[2] = {nullptr, nullptr};
  double *psum = nullptr, *sbar = nullptr, *lpart = nullptr, *lacc = nullptr;
  double *pacc = nullptr, *pred = nullptr;   // parameter-gradient accumulators, sums
  double pgrad[3] = {0, 0, 0};               // d/d(Wi, beta, epsPTT) of the last VJP
  // rollout: checkpoints of U records, the end state, the tape of one interval
  int T = 0;                 // steps of the last forward rollout
  std::vector<double *> ckpt;
  std::vector<sv_ctrl> ckpt_ctrl;
  double *Uend = nullptr;
  sv_ctrl ctrl_end{};
  std::vector<double *> tape;    // [every][8]
  double t_fwd = 0, t_bwd = 0;   // ms of the last rollout / vjp
  int lpn = 1;                   // lanes per node of the transposed gathers (1 or 4)
};

// sub-blocks per 256-node block of the reverse sweeps with lpn lanes per node
static int subl_of(int lpn) { return SUBS * lpn; }

void sv_adj_free(sv_ctx *c) {
  delete c->adj;
  c->adj = nullptr;
}

static int adj_alloc(sv_ctx *c, double **p, size_t n) {
  return sv_alloc(c, (void **)p, sizeof(double) * (n ? n : 1));
}

static int zero_dev(sv_ctx *c, double *p, int64_t n) {
  if (n <= 0) return 0;
  k_zero<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(p, n);
  SV_HIP(c, hipGetLastError());
  return 0;
}

// Transposed stencil of the folded SELL lists: for each j the entries (i, k)
// with scol = j, in ascending i, stage weights zero beyond ij_count_small.
static int build_transpose(sv_ctx *c, sv_adj *A) {
  const sv_loopl &LP = c->L;
  const int n = c->npfb, ns = LP.nslices;
  std::vector<int> cnt(n), cs(n), scol(LP.nnz);
  std::vector<int64_t> off(ns);
  std::vector<double> wx(LP.nnz), wy(LP.nnz), wl(LP.nnz), wh(LP.nnz);
  SV_HIP(c, hipMemcpy(cnt.data(), LP.cnt, sizeof(int) * n, hipMemcpyDeviceToHost));
  SV_HIP(c, hipMemcpy(cs.data(), LP.cnt_small, sizeof(int) * n, hipMemcpyDeviceToHost));
  SV_HIP(c, hipMemcpy(off.data(), LP.off, sizeof(int64_t) * ns, hipMemcpyDeviceToHost));
  const size_t L = sizeof(double) * LP.nnz;
  SV_HIP(c, hipMemcpy(scol.data(), LP.col, sizeof(int) * LP.nnz, hipMemcpyDeviceToHost));
  SV_HIP(c, hipMemcpy(wx.data(), LP.wx, L, hipMemcpyDeviceToHost));
  SV_HIP(c, hipMemcpy(wy.data(), LP.wy, L, hipMemcpyDeviceToHost));
  SV_HIP(c, hipMemcpy(wl.data(), LP.wl, L, hipMemcpyDeviceToHost));
  SV_HIP(c, hipMemcpy(wh.data(), LP.wh, L, hipMemcpyDeviceToHost));
  std::vector<int> tc(n, 0);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < cnt[i]; k++) {
      const int j = scol[off[i >> 6] + (int64_t)k * SV_WAVE + (i & 63)];
      if (j < 0 || j >= n) return sv_fail(c, SV_E_STATE, "stencil column out of range");
      tc[j]++;
    }
  const int nsl = (n + SV_WAVE - 1) / SV_WAVE;
  std::vector<int64_t> toff(nsl + 1);
  int64_t tot = 0;
  for (int s = 0; s < nsl; s++) {
    int w = 0;
    for (int l = 0; l < SV_WAVE && s * SV_WAVE + l < n; l++) w = std::max(w, tc[s * SV_WAVE + l]);
    toff[s] = tot;
    tot += (int64_t)w * SV_WAVE;
  }
  toff[nsl] = tot;
  A->tnnz = tot;
  std::vector<int> tcol(tot);
  std::vector<double> tx(tot, 0.0), ty(tot, 0.0), tl(tot, 0.0), th(tot, 0.0);
  for (int s = 0; s < nsl; s++)       // padding: own node, zero weights
    for (int64_t e = toff[s]; e < toff[s + 1]; e++)
      tcol[e] = std::min(s * SV_WAVE + (int)((e - toff[s]) % SV_WAVE), n - 1);
  std::vector<int> fill(n, 0);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < cnt[i]; k++) {
      const int64_t e = off[i >> 6] + (int64_t)k * SV_WAVE + (i & 63);
      const int j = scol[e];
      const int64_t d = toff[j >> 6] + (int64_t)(fill[j]++) * SV_WAVE + (j & 63);
      tcol[d] = i;
      const bool small = k < cs[i];
      tx[d] = small ? wx[e] : 0.0;
      ty[d] = small ? wy[e] : 0.0;
      tl[d] = small ? wl[e] : 0.0;
      th[d] = wh[e];
    }
  if (sv_alloc(c, (void **)&A->toff, sizeof(int64_t) * (nsl + 1)) ||
      sv_alloc(c, (void **)&A->tcnt, sizeof(int) * n) ||
      sv_alloc(c, (void **)&A->tcol, sizeof(int) * (tot + 1)) || adj_alloc(c, &A->twx, tot) ||
      adj_alloc(c, &A->twy, tot) || adj_alloc(c, &A->twl, tot) || adj_alloc(c, &A->twh, tot))
    return SV_E_NOMEM;
  SV_HIP(c, hipMemcpy(A->toff, toff.data(), sizeof(int64_t) * (nsl + 1), hipMemcpyHostToDevice));
  SV_HIP(c, hipMemcpy(A->tcnt, tc.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  SV_HIP(c, hipMemcpy(A->tcol, tcol.data(), sizeof(int) * tot, hipMemcpyHostToDevice));
  SV_HIP(c, hipMemcpy(A->twx, tx.data(), sizeof(double) * tot, hipMemcpyHostToDevice));
  SV_HIP(c, hipMemcpy(A->twy, ty.data(), sizeof(double) * tot, hipMemcpyHostToDevice));
  SV_HIP(c, hipMemcpy(A->twl, tl.data(), sizeof(double) * tot, hipMemcpyHostToDevice));
  SV_HIP(c, hipMemcpy(A->twh, th.data(), sizeof(double) * tot, hipMemcpyHostToDevice));
  return 0;
}

static int ensure_tape(sv_ctx *c, sv_adj *A) {
  const size_t N = (size_t)c->ntot;
  while ((int)A->tape.size() < 8 * A->every) {
    double *p = nullptr;
    const size_t per = ((A->tape.size() & 1) ? SV_TAPE_D : SV_US) * N;
    if (adj_alloc(c, &p, per)) return SV_E_NOMEM;
    A->tape.push_back(p);
  }
  return 0;
}

// pointwise-adjoint arguments for stage k (1..4) of a tape slot, buffer set q
static PointArgs point_args(sv_ctx *c, sv_adj *A, double *const *tape, int k, int q) {
  return PointArgs{c->blk_seg, c->segk, tape[2 * (k - 1)], tape[2 * (k - 1) + 1], A->Rbar,
                   c->L.gsx, c->L.gsy, c->L.lsum, A->dbar[q], A->dbarm[q], A->own[q],
                   A->pacc};
}

// Sum the accumulators and map d/d(visc_solvent, cp, lambda, eps) to the
// parameters through the derived constants (load_data.F90:132-148):
// visc_solvent = beta*vt, lambda = Wi*L/U, cp = (1 - beta)*vt/lambda.
static int param_grad_finish(sv_ctx *c, sv_adj *A) {
  k_param_reduce<<<1, 1024, 0, c->stream>>>(A->pacc, c->npfb, A->pred);
  SV_HIP(c, hipGetLastError());
  double g[4];
  SV_HIP(c, hipMemcpyAsync(g, A->pred, sizeof(g), hipMemcpyDeviceToHost, c->stream));
  SV_HIP(c, hipStreamSynchronize(c->stream));
  const sv_physk &P = c->pk;
  const double vt = P.visc_total, lam = P.lambda, cp = P.coef_polymeric;
  const double dlam_dWi = P.L_char / P.U_char;
  A->pgrad[0] = g[2] * dlam_dWi - g[1] * cp / lam * dlam_dWi;
  A->pgrad[1] = g[0] * vt - g[1] * vt / lam;
  A->pgrad[2] = g[3];
  return 0;
}

// transposed gather of stage STAGE (its records in buffer set q); for STAGE
// 2..4 fused with the pointwise adjoint of STAGE-1 into buffer set 1-q
template <int STAGE>
static int adj_trans(sv_ctx *c, sv_adj *A, double *const *tape, int q, double seed) {
  const int k = STAGE == 0 ? 1 : STAGE;
  TransArgs t{c->blk_seg, c->segk, A->toff, A->tcnt, A->tcol, A->twx, A->twy, A->twl,
              A->dbar[q], A->dbarm[q], A->own[q], tape[2 * (k - 1)], c->ctrl, A->Ubar, A->Rbar,
              A->regbar, seed, A->psum, PointArgs{}};
  if (STAGE >= 2) t.nxt = point_args(c, A, tape, STAGE - 1, 1 - q);
  const int grid = c->nblk * subl_of(A->lpn);
  if (A->lpn == 4) {
    if (c->any_mdiff) k_adj_trans<STAGE, true, 4><<<grid, AB, 0, c->stream>>>(t);
    else k_adj_trans<STAGE, false, 4><<<grid, AB, 0, c->stream>>>(t);
  } else {
    if (c->any_mdiff) k_adj_trans<STAGE, true, 1><<<grid, AB, 0, c->stream>>>(t);
    else k_adj_trans<STAGE, false, 1><<<grid, AB, 0, c->stream>>>(t);
  }
  SV_HIP(c, hipGetLastError());
  return 0;
}

// calc_all_rhs vector-Jacobian product at tape slot 0 (Rbar in A->Rbar)
static int adj_rhs(sv_ctx *c, sv_adj *A, double *const *tape) {
  const PointArgs p = point_args(c, A, tape, 1, 0);
  if (c->any_mdiff) k_adj_point<true><<<c->nblk * SUBS, AB, 0, c->stream>>>(p);
  else k_adj_point<false><<<c->nblk * SUBS, AB, 0, c->stream>>>(p);
  SV_HIP(c, hipGetLastError());
  return adj_trans<0>(c, A, tape, 0, 0.0);
}

// adjoint of one step (tape slot): Ubar (after the step, with its mass sums
// in psum) -> Ubar (before the step) + seed on cxx, cyy.  Six launches.
static int adj_step(sv_ctx *c, sv_adj *A, double *const *tape, double seed) {
  const int subl = subl_of(A->lpn), grid = c->nblk * subl;
  k_adj_sum_reduce<<<c->nseg, 1024, 0, c->stream>>>(c->segk, A->psum, subl, A->sbar);
  FiltArgs f{c->blk_seg, c->segk, A->toff, A->tcnt, A->tcol, A->twh, c->L.hsum, c->L.vol,
             A->sbar, A->Ubar, c->ctrl, A->regbar, point_args(c, A, tape, 4, 0)};
  if (A->lpn == 4) {
    if (c->any_mdiff) k_adj_filter<true, 4><<<grid, AB, 0, c->stream>>>(f);
    else k_adj_filter<false, 4><<<grid, AB, 0, c->stream>>>(f);
  } else {
    if (c->any_mdiff) k_adj_filter<true, 1><<<grid, AB, 0, c->stream>>>(f);
    else k_adj_filter<false, 1><<<grid, AB, 0, c->stream>>>(f);
  }
  SV_HIP(c, hipGetLastError());
  if (int rc = adj_trans<4>(c, A, tape, 0, 0.0)) return rc;   // + point 3 -> set 1
  if (int rc = adj_trans<3>(c, A, tape, 1, 0.0)) return rc;   // + point 2 -> set 0
  if (int rc = adj_trans<2>(c, A, tape, 0, 0.0)) return rc;   // + point 1 -> set 1
  return adj_trans<1>(c, A, tape, 1, seed);
}

static int loss_accumulate(sv_ctx *c, sv_adj *A) {
  k_loss_part<<<c->nblk, BS, 0, c->stream>>>(c->blk_seg, c->segk, c->U, A->lpart);
  k_loss_acc<<<c->nseg, 1024, 0, c->stream>>>(c->segk, A->lpart, A->lacc);
  SV_HIP(c, hipGetLastError());
  return 0;
}

// caller-order arrays <-> loop-order records
static int upload_records(sv_ctx *c, double *dst, const double *const f[6]) {
  const int n = c->npfb;
  const std::vector<int> &P = c->L.hperm;
  std::vector<double> rec((size_t)SV_US * c->ntot, 0.0);
  for (int q = 0; q < 6; q++)
    if (f && f[q])
      for (int k = 0; k < n; k++) rec[(size_t)SV_US * k + q] = f[q][P[k]];
  SV_HIP(c, hipMemcpy(dst, rec.data(), sizeof(double) * rec.size(), hipMemcpyHostToDevice));
  return 0;
}

static int download_records(sv_ctx *c, const double *src, double *const f[6]) {
  const int n = c->npfb;
  const std::vector<int> &P = c->L.hperm;
  std::vector<double> rec((size_t)SV_US * n);
  SV_HIP(c, hipMemcpy(rec.data(), src, sizeof(double) * rec.size(), hipMemcpyDeviceToHost));
  for (int q = 0; q < 6; q++)
    if (f[q])
      for (int k = 0; k < n; k++) f[q][P[k]] = rec[(size_t)SV_US * k + q];
  return 0;
}

// caller-order device arrays <-> loop-order records, on the device
struct Fields6 {
  const double *f[6];
};
struct OutFields6 {
  double *f[6];
};

__global__ void k_gather_records(int n, const int *perm, Fields6 in, double *rec) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = perm[k];
#pragma unroll
  for (int q = 0; q < 6; q++) rec[(size_t)SV_US * k + q] = in.f[q] ? in.f[q][i] : 0.0;
}

__global__ void k_scatter_records(int n, const int *perm, const double *rec, OutFields6 out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = perm[k];
#pragma unroll
  for (int q = 0; q < 6; q++)
    if (out.f[q]) out.f[q][i] = rec[(size_t)SV_US * k + q];
}

// The eleven output fields of sv_get_state in the caller's node order, field
// after field (out[q * n + i]): ro, rou, rov from the registers, u, v from the
// gather records, psi (LC; left untouched in DI) from the registers, c from
// the gather records.
__global__ void k_get_fields(int n, const int *perm, const double *U, const double *G, int gs,
                             int lc, double *out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = perm[k];
  const double *u = U + (size_t)SV_US * k;
  const double *g = G + (size_t)gs * k;
  out[0 * (size_t)n + i] = u[0];
  out[1 * (size_t)n + i] = u[1];
  out[2 * (size_t)n + i] = u[2];
  out[3 * (size_t)n + i] = g[SV_G_U];
  out[4 * (size_t)n + i] = g[SV_G_V];
  if (lc) {
    out[5 * (size_t)n + i] = u[3];
    out[6 * (size_t)n + i] = u[4];
    out[7 * (size_t)n + i] = u[5];
  }
  out[8 * (size_t)n + i] = g[SV_G_CXX];
  out[9 * (size_t)n + i] = g[SV_G_CXY];
  out[10 * (size_t)n + i] = g[SV_G_CYY];
}

// nf fields of n values from loop order (in[f * n + k]) to the caller's
// order (out[f * n + perm[k]])
__global__ void k_to_caller_order(int n, int nf, const int *perm, const double *in,
                                  double *out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = perm[k];
  for (int f = 0; f < nf; f++) out[(size_t)f * n + i] = in[(size_t)f * n + k];
}

int sv_dev_to_caller_order(sv_ctx *c, int nf, const double *in, double *out) {
  const int n = c->npfb;
  k_to_caller_order<<<(n + 255) / 256, 256, 0, c->stream>>>(n, nf, c->L.perm, in, out);
  SV_HIP(c, hipGetLastError());
  return 0;
}

int sv_dev_get_fields(sv_ctx *c, double *out) {
  const int n = c->npfb;
  k_get_fields<<<(n + 255) / 256, 256, 0, c->stream>>>(n, c->L.perm, c->U, c->GA, c->gs,
                                                       c->cfg.form == SV_FORM_LC, out);
  SV_HIP(c, hipGetLastError());
  return 0;
}

int sv_dev_gather_records(sv_ctx *c, double *dst, const double *const f[6]) {
  Fields6 in{};
  for (int q = 0; q < 6; q++) in.f[q] = f ? f[q] : nullptr;
  const int n = c->npfb;
  k_gather_records<<<(n + 255) / 256, 256, 0, c->stream>>>(n, c->L.perm, in, dst);
  SV_HIP(c, hipGetLastError());
  return 0;
}

int sv_dev_scatter_records(sv_ctx *c, const double *src, double *const f[6]) {
  OutFields6 out{};
  for (int q = 0; q < 6; q++) out.f[q] = f[q];
  const int n = c->npfb;
  k_scatter_records<<<(n + 255) / 256, 256, 0, c->stream>>>(n, c->L.perm, src, out);
  SV_HIP(c, hipGetLastError());
  return 0;
}

#define SV_ADJ_CHECK(c, cond, ...)                            \
  do {                                                        \
    if (!(cond)) return sv_fail((c), SV_E_ARG, __VA_ARGS__);  \
  } while (0)

extern "C" {

int sv_adj_setup(sv_ctx *c, int32_t checkpoint_every) {
  SV_ADJ_CHECK(c, c && c->have_state, "sv_set_initial must come first");
  SV_ADJ_CHECK(c, !c->is_ensemble, "not available on an ensemble context");
  SV_ADJ_CHECK(c, c->cfg.form == SV_FORM_DI, "the adjoint is implemented for the DI form");
  SV_ADJ_CHECK(c, c->nb == 0, "the adjoint is implemented for node sets without boundary nodes");
  SV_ADJ_CHECK(c, !c->has_mirror, "the adjoint is implemented for periodic node sets (no "
               "symmetric/no-slip mirror ghosts)");
  SV_ADJ_CHECK(c, checkpoint_every >= 1 && checkpoint_every <= 1000,
               "checkpoint interval must be in [1, 1000]");
  SV_HIP(c, hipSetDevice(c->device));
  sv_ctrl h{};
  SV_HIP(c, hipMemcpy(&h, c->ctrl, sizeof(h), hipMemcpyDeviceToHost));
  SV_ADJ_CHECK(c, h.dt_mode == SV_DT_FIXED,
               "the adjoint needs a fixed time step (sv_set_control with SV_DT_FIXED)");
  if (!c->adj) {
    sv_adj *A = new sv_adj();
    c->adj = A;
    const size_t R = (size_t)SV_US * c->ntot;
    if (int rc = build_transpose(c, A)) return rc;
    if (adj_alloc(c, &A->Ubar, R) || adj_alloc(c, &A->Rbar, R) || adj_alloc(c, &A->regbar, R) ||
        adj_alloc(c, &A->own[0], R) || adj_alloc(c, &A->own[1], R) ||
        adj_alloc(c, &A->dbar[0], (size_t)SV_TAPE_D * c->ntot) ||
        adj_alloc(c, &A->dbar[1], (size_t)SV_TAPE_D * c->ntot) ||
        adj_alloc(c, &A->dbarm[0], (size_t)4 * c->ntot) ||
        adj_alloc(c, &A->dbarm[1], (size_t)4 * c->ntot) ||
        adj_alloc(c, &A->psum, (size_t)c->nblk * subl_of(4)) ||
        adj_alloc(c, &A->sbar, c->nseg) || adj_alloc(c, &A->lpart, c->nblk) ||
        adj_alloc(c, &A->lacc, c->nseg) || adj_alloc(c, &A->Uend, R) ||
        adj_alloc(c, &A->pacc, (size_t)4 * c->ntot) || adj_alloc(c, &A->pred, 4))
      return SV_E_NOMEM;
  }
  c->adj->every = checkpoint_every;
  // lanes per node of the transposed gathers: 4 when one lane per node would
  // give fewer than 8 waves per CU (e.g. a 128x128 lattice: 256 waves)
  {
    int ncu = 0;
    SV_HIP(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    const int64_t waves = (int64_t)c->nblk * SUBS;
    c->adj->lpn = c->adj_lanes == 1 || c->adj_lanes == 4 ? c->adj_lanes
                  : (waves < 8 * (int64_t)ncu ? 4 : 1);
  }
  return 0;
}

// calc_all_rhs vector-Jacobian product at the current state: ubar = J^T rbar
int sv_rhs_vjp(sv_ctx *c, const double *const rbar[6], double *const ubar[6]) {
  SV_ADJ_CHECK(c, c && c->adj, "sv_adj_setup must come first");
  SV_ADJ_CHECK(c, rbar && ubar, "null argument");
  sv_adj *A = c->adj;
  SV_HIP(c, hipSetDevice(c->device));
  if (int rc = ensure_tape(c, A)) return rc;
  {
    DevTmp<double> scratch;
    SV_HIP(c, scratch.alloc(6 * (size_t)c->npfb));
    if (int rc = sv_dev_rhs(c, 1, scratch, A->tape[0], A->tape[1])) return rc;
  }
  if (int r2 = upload_records(c, A->Rbar, rbar)) return r2;
  if (int r2 = zero_dev(c, A->pacc, (int64_t)4 * c->ntot)) return r2;
  if (int r2 = adj_rhs(c, A, A->tape.data())) return r2;
  if (int r2 = param_grad_finish(c, A)) return r2;
  return download_records(c, A->Ubar, ubar);
}

// nsteps of the fixed-dt loop from the current state, storing checkpoints;
// *loss = (1/nsteps) sum_t mean_i (cxx + cyy - 2) after each step
int sv_rollout_forward(sv_ctx *c, int32_t nsteps, double *loss) {
  SV_ADJ_CHECK(c, c && c->adj, "sv_adj_setup must come first");
  SV_ADJ_CHECK(c, nsteps >= 1, "nsteps must be >= 1");
  sv_adj *A = c->adj;
  SV_HIP(c, hipSetDevice(c->device));
  {
    sv_ctrl h{};
    SV_HIP(c, hipMemcpy(&h, c->ctrl, sizeof(h), hipMemcpyDeviceToHost));
    SV_ADJ_CHECK(c, h.dt_mode == SV_DT_FIXED, "the adjoint needs a fixed time step");
  }
  const int nck = (nsteps + A->every - 1) / A->every;
  const size_t R = (size_t)SV_US * c->ntot;
  while ((int)A->ckpt.size() < nck) {
    double *p = nullptr;
    if (adj_alloc(c, &p, R)) return SV_E_NOMEM;
    A->ckpt.push_back(p);
  }
  A->ckpt_ctrl.assign(nck, sv_ctrl{});
  if (int rc = zero_dev(c, A->lacc, c->nseg)) return rc;
  EventTmp e0, e1;
  SV_HIP(c, e0.create());
  SV_HIP(c, e1.create());
  SV_HIP(c, hipEventRecord(e0, c->stream));
  for (int s = 0; s < nsteps; s++) {
    if (s % A->every == 0) {
      const int k = s / A->every;
      SV_HIP(c, hipMemcpyAsync(A->ckpt[k], c->U, sizeof(double) * R, hipMemcpyDeviceToDevice,
                               c->stream));
      SV_HIP(c, hipMemcpyAsync(&A->ckpt_ctrl[k], c->ctrl, sizeof(sv_ctrl),
                               hipMemcpyDeviceToHost, c->stream));
    }
    if (int rc = sv_dev_step(c, nullptr)) return rc;
    if (int rc = loss_accumulate(c, A)) return rc;
  }
  SV_HIP(c, hipEventRecord(e1, c->stream));
  SV_HIP(c, hipMemcpyAsync(A->Uend, c->U, sizeof(double) * R, hipMemcpyDeviceToDevice,
                           c->stream));
  SV_HIP(c, hipMemcpyAsync(&A->ctrl_end, c->ctrl, sizeof(sv_ctrl), hipMemcpyDeviceToHost,
                           c->stream));
  double acc = 0.0;
  SV_HIP(c, hipMemcpyAsync(&acc, A->lacc, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  SV_HIP(c, hipStreamSynchronize(c->stream));
  float ms = 0;
  SV_HIP(c, hipEventElapsedTime(&ms, e0, e1));
  A->t_fwd = ms;
  A->T = nsteps;
  if (loss) *loss = acc / nsteps;
  return 0;
}

// Gradient of  loss_bar * L + <ubar_final, U_T>  with respect to the state
// (ro, rou, rov, cxx, cxy, cyy) at the start of the last sv_rollout_forward.
// ubar_final may be NULL (or hold NULL fields).  The context's state is left
// at the end of the rollout.
// the loss term's seed on cxx and cyy of every real node's record
__global__ void k_seed_loss(int n, double seed, double *rec) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  rec[(size_t)SV_US * k + 3] += seed;
  rec[(size_t)SV_US * k + 5] += seed;
}

static int rollout_vjp(sv_ctx *c, double loss_bar, const double *const ubar_final[6],
                       double *const grad0[6], bool device_io) {
  SV_ADJ_CHECK(c, c && c->adj && c->adj->T > 0, "sv_rollout_forward must come first");
  SV_ADJ_CHECK(c, grad0, "null output");
  sv_adj *A = c->adj;
  SV_HIP(c, hipSetDevice(c->device));
  if (int rc = ensure_tape(c, A)) return rc;
  const int T = A->T, E = A->every;
  const double seed = loss_bar / ((double)T * c->npfb);
  const size_t R = (size_t)SV_US * c->ntot;
  if (device_io) {
    if (int rc = zero_dev(c, A->Ubar, (int64_t)R)) return rc;
    if (int rc = sv_dev_gather_records(c, A->Ubar, ubar_final)) return rc;
  } else if (int rc = upload_records(c, A->Ubar, ubar_final)) {
    return rc;
  }
  k_seed_loss<<<(c->npfb + 255) / 256, 256, 0, c->stream>>>(c->npfb, seed, A->Ubar);
  SV_HIP(c, hipGetLastError());
  EventTmp e0, e1;
  SV_HIP(c, e0.create());
  SV_HIP(c, e1.create());
  SV_HIP(c, hipEventRecord(e0, c->stream));
  if (int rc = zero_dev(c, A->pacc, (int64_t)4 * c->ntot)) return rc;
  if (A->lpn == 4)
    k_adj_sum<4><<<c->nblk * subl_of(4), AB, 0, c->stream>>>(c->blk_seg, c->segk, A->Ubar,
                                                             A->psum);
  else
    k_adj_sum<1><<<c->nblk * SUBS, AB, 0, c->stream>>>(c->blk_seg, c->segk, A->Ubar, A->psum);
  SV_HIP(c, hipGetLastError());
  const int nck = (T + E - 1) / E;
  for (int k = nck - 1; k >= 0; k--) {
    const int s0 = k * E, s1 = std::min(T, s0 + E);
    // recompute the interval from its checkpoint with the tape on
    SV_HIP(c, hipMemcpyAsync(c->U, A->ckpt[k], sizeof(double) * R, hipMemcpyDeviceToDevice,
                             c->stream));
    SV_HIP(c, hipMemcpyAsync(c->ctrl, &A->ckpt_ctrl[k], sizeof(sv_ctrl), hipMemcpyHostToDevice,
                             c->stream));
    if (int rc = sv_dev_init_state(c)) return rc;
    for (int s = s0; s < s1; s++)
      if (int rc = sv_dev_step_tape(c, A->tape.data() + 8 * (s - s0))) return rc;
    for (int s = s1 - 1; s >= s0; s--)
      if (int rc = adj_step(c, A, A->tape.data() + 8 * (s - s0), s >= 1 ? seed : 0.0))
        return rc;
  }
  SV_HIP(c, hipEventRecord(e1, c->stream));
  SV_HIP(c, hipStreamSynchronize(c->stream));
  float ms = 0;
  SV_HIP(c, hipEventElapsedTime(&ms, e0, e1));
  A->t_bwd = ms;
  if (int rc = param_grad_finish(c, A)) return rc;
  if (device_io) {
    if (int rc = sv_dev_scatter_records(c, A->Ubar, grad0)) return rc;
    SV_HIP(c, hipStreamSynchronize(c->stream));
  } else if (int rc = download_records(c, A->Ubar, grad0)) {
    return rc;
  }
  // leave the state at the end of the rollout
  SV_HIP(c, hipMemcpy(c->U, A->Uend, sizeof(double) * R, hipMemcpyDeviceToDevice));
  SV_HIP(c, hipMemcpy(c->ctrl, &A->ctrl_end, sizeof(sv_ctrl), hipMemcpyHostToDevice));
  return sv_dev_init_state(c);
}

int sv_rollout_vjp(sv_ctx *c, double loss_bar, const double *const ubar_final[6],
                   double *const grad0[6]) {
  return rollout_vjp(c, loss_bar, ubar_final, grad0, false);
}

int sv_rollout_vjp_device(sv_ctx *c, double loss_bar, const double *const ubar_final[6],
                          double *const grad0[6], void *stream) {
  SV_ADJ_CHECK(c, c && grad0, "null argument");
  for (int q = 0; q < 6; q++) {
    if (int rc = sv_check_device_ptr(c, grad0[q], "grad0")) return rc;
    if (ubar_final && ubar_final[q])
      if (int rc = sv_check_device_ptr(c, ubar_final[q], "ubar_final")) return rc;
  }
  // the caller's stream, the null stream included (a framework's default
  // stream): its pending work on the arrays completes first
  SV_HIP(c, hipStreamSynchronize((hipStream_t)stream));
  return rollout_vjp(c, loss_bar, ubar_final, grad0, true);
}

int sv_adj_param_grad(sv_ctx *c, double grad[3]) {
  SV_ADJ_CHECK(c, c && c->adj && grad, "sv_adj_setup must come first");
  for (int k = 0; k < 3; k++) grad[k] = c->adj->pgrad[k];
  return 0;
}

int sv_adj_timing(sv_ctx *c, double *fwd_ms, double *vjp_ms) {
  SV_ADJ_CHECK(c, c && c->adj, "sv_adj_setup must come first");
  if (fwd_ms) *fwd_ms = c->adj->t_fwd;
  if (vjp_ms) *vjp_ms = c->adj->t_bwd;
  return 0;
}

}  // extern "C"

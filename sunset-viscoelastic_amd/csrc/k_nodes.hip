// k_nodes.hip -- ghost (mirror) creation, cell-list neighbour search and the
// SELL-64 layout of the stencils.
//
// Reference behaviour reproduced:
//  * create_mirror_particles (mirror_boundaries.F90:34-261): ghosts within
//    ss*h*2 of periodic/symmetric/no-slip edges, numbered node by node in the
//    order left/right, lower/upper, four corners; positions with the same
//    left-to-right arithmetic (contraction off) so distances are bit-identical.
//  * find_neighbours (neighbours.F90:32-341): every j with
//    |r_i - r_j|^2 <= (h_i*ss)^2 (self included), r^2 = dx*dx + dy*dy without
//    FMA contraction, lists sorted by (r^2, j); ij_count_small counts
//    r^2 <= (h_small_i*ss)^2.
// The cell grid is this library's own: cells of side ss*max(h) over the box
// padded by its size on each side, so a 3x3 block of cells covers every
// stencil whatever h_i is.
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include <climits>
#include <cstdint>
#include <thread>

#include "sv_internal.h"

#define SS 2.0

namespace {

struct GhostDom {
  double xmin, xmax, ymin, ymax;
  int xl, xu, yl, yu;
};

// Enumerate the ghosts of node i in reference order.  emit(gx, gy, vrel).
template <class F>
__device__ void ghost_enum(const GhostDom &g, double xi, double yi, double hi, F emit) {
#pragma clang fp contract(off)
  const double lim = SS * hi * 2.0;
  if (xi <= g.xmin + lim) {
    if (g.xl == 1) emit(xi + g.xmax - g.xmin, yi, 1);
    else if (g.xl == 2) emit(2.0 * g.xmin - xi, yi, 2);
  }
  if (xi >= g.xmax - lim) {
    if (g.xu == 1) emit(xi - g.xmax + g.xmin, yi, 1);
    else if (g.xu == 2) emit(2.0 * g.xmax - xi, yi, 2);
  }
  if (yi <= g.ymin + lim) {
    if (g.yl == 1) emit(xi, yi + g.ymax - g.ymin, 1);
    else if (g.yl == 2) emit(xi, 2.0 * g.ymin - yi, 3);
    else if (g.yl == 3) emit(xi, 2.0 * g.ymin - yi, 4);
  }
  if (yi >= g.ymax - lim) {
    if (g.yu == 1) emit(xi, yi - g.ymax + g.ymin, 1);
    else if (g.yu == 2) emit(xi, 2.0 * g.ymax - yi, 3);
    else if (g.yu == 3) emit(xi, 2.0 * g.ymax - yi, 4);
  }
  const double cx[4] = {g.xmin, g.xmax, g.xmin, g.xmax};
  const double cy[4] = {g.ymin, g.ymin, g.ymax, g.ymax};
  const int bx[4] = {g.xl, g.xu, g.xl, g.xu};
  const int by[4] = {g.yl, g.yl, g.yu, g.yu};
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const double dx = cx[c] - xi, dy = cy[c] - yi;
    const double cd = sqrt(dx * dx + dy * dy + 0.0 * 0.0);
    if (cd <= lim && bx[c] != 0 && by[c] != 0) {
      const bool perx = bx[c] == 1, pery = by[c] == 1;
      const bool symx = bx[c] == 2, symy = by[c] == 2;
      double px, py;
      if (c == 0 || c == 2) px = perx ? xi + g.xmax - g.xmin : 2.0 * g.xmin - xi;
      else px = perx ? xi - g.xmax + g.xmin : 2.0 * g.xmax - xi;
      if (c == 0 || c == 1) py = pery ? yi + g.ymax - g.ymin : 2.0 * g.ymin - yi;
      else py = pery ? yi - g.ymax + g.ymin : 2.0 * g.ymax - yi;
      int vr = 0;
      if (perx && pery) vr = 1;
      else if (symx && pery) vr = 2;
      else if (perx && symy) vr = 3;
      else if (symx && symy) vr = 4;
      if (vr) emit(px, py, vr);
      else emit(0.0, 0.0, 0);   // the reference counts the ghost but leaves rp = 0
    }
  }
}

__global__ void k_ghost_count(int npfb, GhostDom g, const double *x, const double *y,
                              const double *h, int *gcount) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npfb) return;
  int n = 0;
  ghost_enum(g, x[i], y[i], h[i], [&](double, double, int) { n++; });
  gcount[i] = n;
}

__global__ void k_ghost_write(int npfb, GhostDom g, double *x, double *y, double *h, double *hs,
                              double *s, int *nt, int *irel, int *vrel, const int *goff) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npfb) return;
  int k = npfb + goff[i];
  const double hi = h[i], hsi = hs[i], si = s[i];
  const int nti = nt[i];
  ghost_enum(g, x[i], y[i], hi, [&](double gx, double gy, int vr) {
    x[k] = gx; y[k] = gy; irel[k] = i; vrel[k] = vr;
    h[k] = hi; hs[k] = hsi; s[k] = si; nt[k] = nti;
    k++;
  });
  irel[i] = i;
  vrel[i] = 0;
}

struct Grid {
  double x0, y0, inv;
  int ncx, ncy;
  // the reference's cells (cell_generator, neighbours.F90:75-93): size
  // ss*h0 over the domain padded by its size on every side; a node sees the
  // 3x3 cells around its own (neighbour_cells, :95-157), which cuts the disc
  // of a node with h > h0
  double rx0, ry0, rinv, h0;
};

__device__ __forceinline__ int cell_of(const Grid &G, double x, double y, int *cx, int *cy) {
  int ix = (int)((x - G.x0) * G.inv);
  int iy = (int)((y - G.y0) * G.inv);
  ix = ix < 0 ? 0 : (ix >= G.ncx ? G.ncx - 1 : ix);
  iy = iy < 0 ? 0 : (iy >= G.ncy ? G.ncy - 1 : iy);
  *cx = ix;
  *cy = iy;
  return iy * G.ncx + ix;
}

__global__ void k_cell_count(int np, Grid G, const double *x, const double *y, int *cellof,
                             int *ccount) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  int cx, cy;
  const int c = cell_of(G, x[k], y[k], &cx, &cy);
  cellof[k] = c;
  atomicAdd(&ccount[c], 1);
}

__global__ void k_cell_fill(int np, const int *cellof, const int *cstart, int *cfill, int *ip) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const int c = cellof[k];
  const int pos = atomicAdd(&cfill[c], 1);
  ip[cstart[c] + pos] = k;
}

__device__ __forceinline__ double dist2(double xi, double yi, double xj, double yj) {
#pragma clang fp contract(off)
  const double dx = xi - xj, dy = yi - yj;
  double t = dx * dx;
  t = t + dy * dy;
  return t;
}

// Wave-scope ordering of LDS traffic between the lanes of one wavefront.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int NW = 4;   // nodes (wavefronts) per block of the neighbour search

// One wavefront per node.  The cells within the node's own search radius
// 2 h_i (one cell of margin) are scanned row by row: a row of cells is one
// contiguous range of the cell-sorted node list, taken 64 candidates at a
// time, and matches are compacted with a ballot.  Pass 0 counts; pass 1
// collects (j, r^2) in LDS, ranks every entry by (r^2, j) -- the order of the
// reference's sorted lists (neighbours.F90:247-341) -- and writes the row in
// that order, with the count of entries inside 2 h_small.  The cell size is
// 2 h_min, so nodes of every size scan a few candidates per neighbour.  A node
// with h > h0 (boundary nodes at order 4: hovs_bound 2.1 > hovs 1.8,
// common_parameter.F90:46-47) keeps, as the reference does, only the
// candidates inside the 3x3 reference cells of size 2 h0 around it.
template <int PASS>
__global__ __launch_bounds__(NW * SV_WAVE) void k_neigh(
    int npfb, Grid G, const double *x, const double *y, const double *h, const double *hs,
    const int *cstart, const int *ip, int *cnt, int *cnt_small, const int64_t *row_ptr,
    int *col, int cap) {
#pragma clang fp contract(off)
  extern __shared__ double nb_lds[];
  const int wave = (int)threadIdx.x / SV_WAVE, lane = (int)threadIdx.x % SV_WAVE;
  const int i = (int)blockIdx.x * NW + wave;
  if (i >= npfb) return;   // whole wavefronts; no block-wide barrier follows
  double *r2s = nb_lds + (size_t)wave * cap;
  int *js = reinterpret_cast<int *>(nb_lds + (size_t)NW * cap) + (size_t)wave * cap;
  const double xi = x[i], yi = y[i];
  const double R = h[i] * SS;
  const double st2 = R * R;
  // only a node with h > h0 reaches past the reference's 3x3 cells
  const bool clip = h[i] > G.h0;
  const int rcx = (int)((xi - G.rx0) * G.rinv), rcy = (int)((yi - G.ry0) * G.rinv);
  int ix0 = (int)((xi - R - G.x0) * G.inv) - 1, ix1 = (int)((xi + R - G.x0) * G.inv) + 1;
  int iy0 = (int)((yi - R - G.y0) * G.inv) - 1, iy1 = (int)((yi + R - G.y0) * G.inv) + 1;
  ix0 = ix0 < 0 ? 0 : ix0;
  iy0 = iy0 < 0 ? 0 : iy0;
  ix1 = ix1 >= G.ncx ? G.ncx - 1 : ix1;
  iy1 = iy1 >= G.ncy ? G.ncy - 1 : iy1;
  int n = 0;
  for (int cy = iy0; cy <= iy1; cy++) {
    const int lo = cstart[cy * G.ncx + ix0], hi = cstart[cy * G.ncx + ix1 + 1];
    for (int b0 = lo; b0 < hi; b0 += SV_WAVE) {
      const int jj = b0 + lane;
      bool m = false;
      double r2 = 0.0;
      int j = 0;
      if (jj < hi) {
        j = ip[jj];
        r2 = dist2(xi, yi, x[j], y[j]);
        m = r2 <= st2;
        if (clip && m) {
          const int dcx = (int)((x[j] - G.rx0) * G.rinv) - rcx;
          const int dcy = (int)((y[j] - G.ry0) * G.rinv) - rcy;
          m = dcx >= -1 && dcx <= 1 && dcy >= -1 && dcy <= 1;
        }
      }
      const unsigned long long bal = __ballot(m);
      if (PASS == 1 && m) {
        const int pos = n + __popcll(bal & ((1ull << lane) - 1ull));
        if (pos < cap) {
          r2s[pos] = r2;
          js[pos] = j;
        }
      }
      n += __popcll(bal);
    }
  }
  if (PASS == 0) {
    if (lane == 0) cnt[i] = n;
    return;
  }
  wave_sync();
  // the same scan as pass 0, so n == cnt[i] <= cap; clamped all the same so
  // that a row can never spill into the next one
  n = n < cnt[i] ? n : cnt[i];
  const int64_t base = row_ptr[i];
  const double hsm = hs[i] * SS;
  const double st2s = hsm * hsm;
  int ks = 0;
  for (int a = lane; a < n; a += SV_WAVE) {
    const double ra = r2s[a];
    const int ja = js[a];
    int rank = 0;
    for (int b = 0; b < n; b++) {
      const double rb = r2s[b];
      const int jb = js[b];
      rank += (rb < ra || (rb == ra && jb < ja)) ? 1 : 0;
    }
    col[base + rank] = ja;
    ks += (ra <= st2s) ? 1 : 0;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) ks += __shfl_xor(ks, o);
  if (lane == 0) cnt_small[i] = ks;
}

__global__ void k_sell_pack(int npfb, const int64_t *sell_off, const int *sell_w,
                            const int64_t *row_ptr, const int *cnt, const int *col_raw,
                            const int *irel, const double *wx, const double *wy,
                            const double *wl, const double *wh, int *scol, int *scol_raw,
                            double *swx, double *swy, double *swl, double *swh) {
  // one block of 64 threads per (slice, link-chunk of 4)
  const int s = blockIdx.x;
  const int l = threadIdx.x;
  const int i = s * SV_WAVE + l;
  const int W = sell_w[s];
  const int64_t off = sell_off[s];
  const int k0 = blockIdx.y * 4;
  for (int k = k0; k < W && k < k0 + 4; k++) {
    const int64_t d = off + (int64_t)k * SV_WAVE + l;
    if (i < npfb && k < cnt[i]) {
      const int64_t sidx = row_ptr[i] + k;
      const int j = col_raw[sidx];
      scol_raw[d] = j;
      scol[d] = irel[j];
      swx[d] = wx ? wx[sidx] : 0.0;
      swy[d] = wy ? wy[sidx] : 0.0;
      swl[d] = wl ? wl[sidx] : 0.0;
      swh[d] = wh ? wh[sidx] : 0.0;
    } else {
      const int self = i < npfb ? i : 0;
      scol_raw[d] = self;
      scol[d] = self;
      swx[d] = 0.0;
      swy[d] = 0.0;
      swl[d] = 0.0;
      swh[d] = 0.0;
    }
  }
}

}  // namespace

static int scan_int(sv_ctx *c, const int *in, int *out, int n) {
  size_t tb = 0;
  SV_HIP(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, n, c->stream));
  DevTmp<char> tmp;
  SV_HIP(c, tmp.alloc(tb > 0 ? tb : 16));
  SV_HIP(c, hipcub::DeviceScan::ExclusiveSum((void *)tmp, tb, in, out, n, c->stream));
  SV_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

int sv_dev_create_ghosts(sv_ctx *c) {
  const int n = c->npfb;
  GhostDom g{c->xmin, c->xmax, c->ymin, c->ymax, c->bc[0], c->bc[1], c->bc[2], c->bc[3]};
  DevTmp<int> gcount, goff;
  SV_HIP(c, gcount.alloc(n + 1));
  SV_HIP(c, goff.alloc(n + 1));
  SV_HIP(c, hipMemsetAsync(gcount, 0, sizeof(int) * (n + 1), c->stream));
  const int B = 256, nb = (n + B - 1) / B;
  k_ghost_count<<<nb, B, 0, c->stream>>>(n, g, c->x, c->y, c->h, gcount);
  SV_HIP(c, hipGetLastError());
  if (int rc = scan_int(c, gcount, goff, n + 1)) return rc;
  int ng = 0;
  SV_HIP(c, hipMemcpy(&ng, goff + n, sizeof(int), hipMemcpyDeviceToHost));
  if (n + ng > c->cap) {
    return sv_fail(c, SV_E_ARG, "ghost count %d exceeds node capacity %d", ng, c->cap - n);
  }
  k_ghost_write<<<nb, B, 0, c->stream>>>(n, g, c->x, c->y, c->h, c->hs, c->s, c->node_type,
                                         c->irel, c->vrel, goff);
  SV_HIP(c, hipGetLastError());
  SV_HIP(c, hipStreamSynchronize(c->stream));
  // symmetric (2) or no-slip (3) edges make mirror ghosts with sign flips
  c->has_mirror = false;
  for (int k = 0; k < 4; k++) c->has_mirror |= (c->bc[k] == 2 || c->bc[k] == 3) && ng > 0;
  c->np = n + ng;
  return 0;
}

int sv_dev_find_neighbours(sv_ctx *c) {
  EventTmp e0, e1;
  SV_HIP(c, e0.create());
  SV_HIP(c, e1.create());
  SV_HIP(c, hipEventRecord(e0, c->stream));
  const int np = c->np, npfb = c->npfb;
  const double Lx = c->xmax - c->xmin, Ly = c->ymax - c->ymin;
  double csz = SS * (c->hmin > 0.0 ? c->hmin : c->hmax);
  // keep the padded grid below ~64M cells
  const double area = 9.0 * Lx * Ly;
  if (area / (csz * csz) > 6.4e7) csz = sqrt(area / 6.4e7);
  Grid G;
  G.x0 = c->xmin - Lx;
  G.y0 = c->ymin - Ly;
  G.inv = 1.0 / csz;
  G.ncx = (int)(3.0 * Lx / csz) + 1;
  G.ncy = (int)(3.0 * Ly / csz) + 1;
  G.rx0 = c->xmin - Lx;
  G.ry0 = c->ymin - Ly;
  G.rinv = 1.0 / (1.0 * (SS * c->h0));
  G.h0 = c->h0;
  const int nct = G.ncx * G.ncy;
  DevTmp<int> cellof, ccount, cstart, ip;
  SV_HIP(c, cellof.alloc(np));
  SV_HIP(c, ccount.alloc(nct + 1));
  SV_HIP(c, cstart.alloc(nct + 1));
  SV_HIP(c, ip.alloc(np));
  SV_HIP(c, hipMemsetAsync(ccount, 0, sizeof(int) * (nct + 1), c->stream));
  const int B = 256;
  k_cell_count<<<(np + B - 1) / B, B, 0, c->stream>>>(np, G, c->x, c->y, cellof, ccount);
  SV_HIP(c, hipGetLastError());
  if (int rc = scan_int(c, ccount, cstart, nct + 1)) return rc;
  SV_HIP(c, hipMemsetAsync(ccount, 0, sizeof(int) * (nct + 1), c->stream));
  k_cell_fill<<<(np + B - 1) / B, B, 0, c->stream>>>(np, cellof, cstart, ccount, ip);
  SV_HIP(c, hipGetLastError());
  // counts
  if (!c->cnt) {
    if (sv_alloc(c, (void **)&c->cnt, sizeof(int) * npfb)) return SV_E_NOMEM;
    if (sv_alloc(c, (void **)&c->cnt_small, sizeof(int) * npfb)) return SV_E_NOMEM;
    if (sv_alloc(c, (void **)&c->row_ptr, sizeof(int64_t) * (npfb + 1))) return SV_E_NOMEM;
  }
  const int nb = (npfb + NW - 1) / NW;
  k_neigh<0><<<nb, NW * SV_WAVE, 0, c->stream>>>(npfb, G, c->x, c->y, c->h, c->hs, cstart, ip,
                                                 c->cnt, c->cnt_small, nullptr, nullptr, 0);
  SV_HIP(c, hipGetLastError());
  std::vector<int> hc(npfb);
  SV_HIP(c, hipMemcpyAsync(hc.data(), c->cnt, sizeof(int) * npfb, hipMemcpyDeviceToHost,
                           c->stream));
  SV_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<int64_t> rp(npfb + 1);
  rp[0] = 0;
  int maxk = 0;
  for (int i = 0; i < npfb; i++) {
    rp[i + 1] = rp[i] + hc[i];
    maxk = hc[i] > maxk ? hc[i] : maxk;
  }
  c->maxK = maxk;
  c->nnz = rp[npfb];
  SV_HIP(c, hipMemcpyAsync(c->row_ptr, rp.data(), sizeof(int64_t) * (npfb + 1),
                           hipMemcpyHostToDevice, c->stream));
  if (c->col_raw) sv_free(c, c->col_raw);
  if (sv_alloc(c, (void **)&c->col_raw, sizeof(int) * (c->nnz + 1))) return SV_E_NOMEM;
  const int cap = maxk > 0 ? maxk : 1;
  const size_t shm = (size_t)NW * cap * (sizeof(double) + sizeof(int));
  if (shm > 64 * 1024)
    return sv_fail(c, SV_E_OVERFLOW, "neighbour list of %d entries too long for the search", maxk);
  k_neigh<1><<<nb, NW * SV_WAVE, shm, c->stream>>>(npfb, G, c->x, c->y, c->h, c->hs, cstart, ip,
                                                   c->cnt, c->cnt_small, c->row_ptr, c->col_raw,
                                                   cap);
  SV_HIP(c, hipGetLastError());
  SV_HIP(c, hipEventRecord(e1, c->stream));
  SV_HIP(c, hipStreamSynchronize(c->stream));
  float ms = 0;
  SV_HIP(c, hipEventElapsedTime(&ms, e0, e1));
  c->t_neigh = ms;
  c->have_lists = true;
  c->sell_dirty = true;
  c->L.valid = false;
  return 0;
}

int sv_dev_build_sell(sv_ctx *c) {
  const int npfb = c->npfb;
  const int ns = (npfb + SV_WAVE - 1) / SV_WAVE;
  std::vector<int> hc(npfb);
  SV_HIP(c, hipMemcpy(hc.data(), c->cnt, sizeof(int) * npfb, hipMemcpyDeviceToHost));
  std::vector<int> w(ns);
  std::vector<int64_t> off(ns + 1);
  off[0] = 0;
  int wmax = 0;
  for (int s = 0; s < ns; s++) {
    int m = 0;
    for (int l = 0; l < SV_WAVE && s * SV_WAVE + l < npfb; l++)
      m = hc[s * SV_WAVE + l] > m ? hc[s * SV_WAVE + l] : m;
    w[s] = m;
    wmax = m > wmax ? m : wmax;
    off[s + 1] = off[s] + (int64_t)m * SV_WAVE;
  }
  if (c->nslices != ns || c->sell_nnz != off[ns]) {
    void *bufs[] = {c->sell_off, c->sell_w, c->scol, c->scol_raw, c->swx, c->swy, c->swl, c->swh};
    for (void *b : bufs)
      if (b) sv_free(c, b);
    c->nslices = ns;
    c->sell_nnz = off[ns];
    const size_t L = (size_t)c->sell_nnz + 1;
    if (sv_alloc(c, (void **)&c->sell_off, sizeof(int64_t) * (ns + 1)) ||
        sv_alloc(c, (void **)&c->sell_w, sizeof(int) * (ns + 1)) ||
        sv_alloc(c, (void **)&c->scol, sizeof(int) * L) ||
        sv_alloc(c, (void **)&c->scol_raw, sizeof(int) * L) ||
        sv_alloc(c, (void **)&c->swx, sizeof(double) * L) ||
        sv_alloc(c, (void **)&c->swy, sizeof(double) * L) ||
        sv_alloc(c, (void **)&c->swl, sizeof(double) * L) ||
        sv_alloc(c, (void **)&c->swh, sizeof(double) * L))
      return SV_E_NOMEM;
  }
  SV_HIP(c, hipMemcpy(c->sell_off, off.data(), sizeof(int64_t) * (ns + 1),
                      hipMemcpyHostToDevice));
  SV_HIP(c, hipMemcpy(c->sell_w, w.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
  dim3 grid(ns, (wmax + 3) / 4);
  k_sell_pack<<<grid, SV_WAVE, 0, c->stream>>>(npfb, c->sell_off, c->sell_w, c->row_ptr, c->cnt,
                                               c->col_raw, c->irel, c->wx, c->wy, c->wl, c->wh,
                                               c->scol, c->scol_raw, c->swx, c->swy, c->swl,
                                               c->swh);
  SV_HIP(c, hipGetLastError());
  SV_HIP(c, hipStreamSynchronize(c->stream));
  c->sell_dirty = false;
  return 0;
}

// ---- time-loop layout (Morton order) --------------------------------------
namespace {

// Loop-order SELL packing: block (s, y) fills links [16y, 16y+16) of the 64
// nodes of slice s.  The CSR rows are read 16 links at a time (one cache
// line of each array per row), transposed through LDS, and written as the
// slice's link-major columns (64 consecutive slots per link).  Links past a
// node's count, and slots past npfb, get the padding entry: the node's own
// slot (0 past npfb) with zero weights.
constexpr int PK_K = 16;           // links per block
constexpr int PK_T = 256;          // threads per block
constexpr int PK_LD = PK_K + 1;    // LDS row stride of the tiles
__global__ __launch_bounds__(PK_T) void k_loop_pack(
    int npfb, const int64_t *off, const int *perm, const int *iperm, const int64_t *row_ptr,
    const int *cnt, const int *col_raw, const int *irel, const int *vrel, const double *wx,
    const double *wy, const double *wl, const double *wh, int64_t nnz_max, int *lcol,
    double *lwx, double *lwy, double *lwl, double *lwh, uint8_t *lflip) {
  __shared__ double tx[SV_WAVE * PK_LD], ty[SV_WAVE * PK_LD], tl[SV_WAVE * PK_LD],
      th[SV_WAVE * PK_LD];
  __shared__ int tc[SV_WAVE * PK_LD];
  __shared__ uint8_t tf[SV_WAVE * PK_LD];
  __shared__ int rnode[SV_WAVE], rcnt[SV_WAVE];
  __shared__ int64_t rbase[SV_WAVE];
  const int s = blockIdx.x;
  const int t = threadIdx.x;
  const int64_t o = off[s];
  const int W = (int)((off[s + 1] - o) / SV_WAVE);
  const int k0 = blockIdx.y * PK_K;
  if (k0 >= W) return;
  if (t < SV_WAVE) {
    const int slot = s * SV_WAVE + t;
    const int i = slot < npfb ? perm[slot] : -1;
    rnode[t] = i;
    rcnt[t] = i >= 0 ? cnt[i] : 0;
    rbase[t] = i >= 0 ? row_ptr[i] : 0;
  }
  __syncthreads();
  // gather: 16 consecutive threads read 16 consecutive links of one row
  for (int p = t; p < SV_WAVE * PK_K; p += PK_T) {
    const int r = p / PK_K, kk = p % PK_K, k = k0 + kk;
    const int slot = s * SV_WAVE + r;
    const int e = r * PK_LD + kk;
    if (rnode[r] >= 0 && k < rcnt[r]) {
      const int64_t sidx = rbase[r] + k;
      const int raw = col_raw[sidx];
      tc[e] = iperm[irel[raw]];
      tx[e] = wx[sidx];
      ty[e] = wy[sidx];
      tl[e] = wl[sidx];
      th[e] = wh[sidx];
      const int vr = (lflip && raw >= npfb) ? vrel[raw] : 0;
      tf[e] = (uint8_t)(vr >= 2 ? vr - 1 : 0);
    } else {
      tc[e] = slot < npfb ? slot : 0;
      tx[e] = 0.0;
      ty[e] = 0.0;
      tl[e] = 0.0;
      th[e] = 0.0;
      tf[e] = 0;
    }
  }
  __syncthreads();
  // scatter: 64 consecutive threads write one link's column of the slice
  for (int p = t; p < SV_WAVE * PK_K; p += PK_T) {
    const int kk = p / SV_WAVE, l = p % SV_WAVE, k = k0 + kk;
    if (k >= W) break;
    const int64_t d = o + (int64_t)k * SV_WAVE + l;
    if (d >= nnz_max) break;
    const int e = l * PK_LD + kk;
    lcol[d] = tc[e];
    lwx[d] = tx[e];
    lwy[d] = ty[e];
    lwl[d] = tl[e];
    lwh[d] = th[e];
    if (lflip) lflip[d] = tf[e];
  }
}

template <typename T>
__global__ void k_permute(int n, const int *perm, const T *src, T *dst) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = src[perm[k]];
}

// interleave the low 16 bits of a and b (Morton key)
uint32_t morton2(uint32_t a, uint32_t b) {
  auto spread = [](uint32_t v) {
    v &= 0xffff;
    v = (v | (v << 8)) & 0x00ff00ff;
    v = (v | (v << 4)) & 0x0f0f0f0f;
    v = (v | (v << 2)) & 0x33333333;
    v = (v | (v << 1)) & 0x55555555;
    return v;
  };
  return spread(a) | (spread(b) << 1);
}

}  // namespace


// Per-block halos of the loop layout (see sv_loopl): host pass over the SELL
// columns, O(nnz log K) at setup.
static int build_halos(sv_ctx *c) {
  sv_loopl &L = c->L;
  const int n = c->npfb, ns = L.nslices;
  const int nb = (n + SV_BLOCK - 1) / SV_BLOCK;
  std::vector<int64_t> off(ns + 1);
  std::vector<int> col(L.nnz), cnt(n);
  SV_HIP(c, hipMemcpy(off.data(), L.off, sizeof(int64_t) * (ns + 1), hipMemcpyDeviceToHost));
  SV_HIP(c, hipMemcpy(col.data(), L.col, sizeof(int) * L.nnz, hipMemcpyDeviceToHost));
  SV_HIP(c, hipMemcpy(cnt.data(), L.cnt, sizeof(int) * n, hipMemcpyDeviceToHost));
  if (L.lflip) {
    // a mirror link and a plain link to the same parent are different halo
    // records (the mirror one carries the sign flips)
    std::vector<uint8_t> fl(L.nnz);
    SV_HIP(c, hipMemcpy(fl.data(), L.lflip, L.nnz, hipMemcpyDeviceToHost));
    for (int64_t e = 0; e < L.nnz; e++) col[e] |= (int)fl[e] << SV_FLIP_SHIFT;
  }
  // blocks are independent: each one's halo (sorted unique link targets) and
  // its links' block-local indices, on up to 16 host threads over
  // contiguous block ranges; the halos are concatenated in block order after
  std::vector<std::vector<int>> hb(nb);
  std::vector<uint16_t> lcol(L.nnz, 0);
  constexpr int SPB = SV_BLOCK / SV_WAVE;   // slices per block
  // one block: its halo (the sorted unique link targets) and each link's
  // rank in it.  Targets within a narrow range (the usual case in Morton
  // order) are ranked with a dense marker array in one pass; a wide range
  // (a Morton seam, or flipped mirror keys) takes a sort and binary searches.
  // Both give the same halo and ranks.
  auto one_block = [&](int b, std::vector<int> &h, std::vector<int> &mark) {
    h.clear();
    int lo = INT32_MAX, hi = INT32_MIN;
    for (int s = b * SPB; s < std::min(ns, (b + 1) * SPB); s++)
      for (int l = 0; l < SV_WAVE && s * SV_WAVE + l < n; l++) {
        const int slot = s * SV_WAVE + l;
        for (int k = 0; k < cnt[slot]; k++) {
          const int v = col[off[s] + (int64_t)k * SV_WAVE + l];
          h.push_back(v);
          lo = std::min(lo, v);
          hi = std::max(hi, v);
        }
      }
    if (h.empty()) return;
    const int64_t range = (int64_t)hi - lo + 1;
    const bool dense = range <= 4 * (int64_t)h.size() + 4096;
    if (dense) {
      mark.assign((size_t)range, -1);
      for (int v : h) mark[v - lo] = 0;
      h.clear();
      for (int64_t q = 0; q < range; q++)
        if (mark[q] == 0) {
          mark[q] = (int)h.size();
          h.push_back(lo + (int)q);
        }
    } else {
      std::sort(h.begin(), h.end());
      h.erase(std::unique(h.begin(), h.end()), h.end());
    }
    if (h.size() > 65535) return;   // too large for 16-bit indices: no halo sweeps
    for (int s = b * SPB; s < std::min(ns, (b + 1) * SPB); s++)
      for (int l = 0; l < SV_WAVE && s * SV_WAVE + l < n; l++) {
        const int slot = s * SV_WAVE + l;
        for (int k = 0; k < cnt[slot]; k++) {
          const int64_t e = off[s] + (int64_t)k * SV_WAVE + l;
          lcol[e] = dense ? (uint16_t)mark[col[e] - lo]
                          : (uint16_t)(std::lower_bound(h.begin(), h.end(), col[e]) - h.begin());
        }
      }
  };
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nt = (int)std::min<unsigned>(std::min(hw, 16u), (unsigned)std::max(nb, 1));
  {
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
      pool.emplace_back([&, t] {
        std::vector<int> mark;
        for (int b = (int)((int64_t)nb * t / nt); b < (int)((int64_t)nb * (t + 1) / nt); b++)
          one_block(b, hb[b], mark);
      });
    for (auto &th : pool) th.join();
  }
  std::vector<int64_t> hoff(nb + 1, 0);
  int hmax = 0;
  bool fits = true;
  for (int b = 0; b < nb; b++) {
    hmax = std::max(hmax, (int)hb[b].size());
    if (hb[b].size() > 65535) fits = false;
    hoff[b + 1] = hoff[b] + (int64_t)hb[b].size();
  }
  std::vector<int> hidx((size_t)hoff[nb]);
  for (int b = 0; b < nb; b++) std::copy(hb[b].begin(), hb[b].end(), hidx.begin() + hoff[b]);
  L.hmax = hmax;
  // the LC sweep keeps 72 bytes per halo record in LDS
  L.halo = c->halo_sweeps && fits && (size_t)hmax * 72 <= SV_HALO_LDS_MAX;
  if (!L.halo) return 0;
  if (sv_alloc(c, (void **)&L.hoff, sizeof(int64_t) * (nb + 1)) ||
      sv_alloc(c, (void **)&L.hidx, sizeof(int) * (hidx.size() + 1)) ||
      sv_alloc(c, (void **)&L.lcol, sizeof(uint16_t) * (L.nnz + 1)))
    return SV_E_NOMEM;
  SV_HIP(c, hipMemcpy(L.hoff, hoff.data(), sizeof(int64_t) * (nb + 1), hipMemcpyHostToDevice));
  SV_HIP(c, hipMemcpy(L.hidx, hidx.data(), sizeof(int) * hidx.size(), hipMemcpyHostToDevice));
  SV_HIP(c, hipMemcpy(L.lcol, lcol.data(), sizeof(uint16_t) * L.nnz, hipMemcpyHostToDevice));
  return 0;
}

int sv_dev_build_loop(sv_ctx *c) {
  const int n = c->npfb;
  sv_loopl &L = c->L;
  // permutation
  std::vector<double> x(n), y(n);
  SV_HIP(c, hipMemcpy(x.data(), c->x, sizeof(double) * n, hipMemcpyDeviceToHost));
  SV_HIP(c, hipMemcpy(y.data(), c->y, sizeof(double) * n, hipMemcpyDeviceToHost));
  L.hperm.resize(n);
  for (int i = 0; i < n; i++) L.hperm[i] = i;
  if (c->loop_order) {
    const double sx = 65535.0 / (c->xmax - c->xmin), sy = 65535.0 / (c->ymax - c->ymin);
    std::vector<uint32_t> key(n);
    for (int i = 0; i < n; i++) {
      const double fx = std::min(std::max((x[i] - c->xmin) * sx, 0.0), 65535.0);
      const double fy = std::min(std::max((y[i] - c->ymin) * sy, 0.0), 65535.0);
      key[i] = morton2((uint32_t)fx, (uint32_t)fy);
    }
    // stable LSD radix sort of the 32-bit keys, two 16-bit digits: the same
    // permutation as a stable comparison sort
    std::vector<int> tmp(n);
    std::vector<int> cnt16(65537);
    for (int shift = 0; shift < 32; shift += 16) {
      std::fill(cnt16.begin(), cnt16.end(), 0);
      for (int i = 0; i < n; i++) cnt16[((key[L.hperm[i]] >> shift) & 0xffff) + 1]++;
      for (int d = 1; d <= 65536; d++) cnt16[d] += cnt16[d - 1];
      for (int i = 0; i < n; i++) tmp[cnt16[(key[L.hperm[i]] >> shift) & 0xffff]++] = L.hperm[i];
      L.hperm.swap(tmp);
    }
  }
  std::vector<int> iperm(c->np);
  for (int k = 0; k < n; k++) iperm[L.hperm[k]] = k;
  for (int j = n; j < c->np; j++) iperm[j] = -1;   // ghosts are folded before lookup
  std::vector<int> cnt(n);
  SV_HIP(c, hipMemcpy(cnt.data(), c->cnt, sizeof(int) * n, hipMemcpyDeviceToHost));
  const int ns = (n + SV_WAVE - 1) / SV_WAVE;
  std::vector<int64_t> off(ns + 1);
  off[0] = 0;
  int wmax = 0;
  for (int s = 0; s < ns; s++) {
    int m = 0;
    for (int l = 0; l < SV_WAVE && s * SV_WAVE + l < n; l++) m = std::max(m, cnt[L.hperm[s * SV_WAVE + l]]);
    wmax = std::max(wmax, m);
    off[s + 1] = off[s] + (int64_t)m * SV_WAVE;
  }
  // (re)allocate
  void *old[] = {L.off, L.col, L.wx, L.wy, L.wl, L.wh, L.cnt, L.cnt_small, L.gsx, L.gsy,
                 L.lsum, L.hsum, L.vol, L.s, L.perm, L.iperm, L.hoff, L.hidx, L.lcol,
                 L.bidx, L.brec, L.lflip};
  for (void *b : old)
    if (b) sv_free(c, b);
  std::vector<int> hp = std::move(L.hperm);
  L = sv_loopl{};
  L.nslices = ns;
  L.nnz = off[ns];
  L.hperm = std::move(hp);
  const size_t E = (size_t)L.nnz + 1, D = sizeof(double), I = sizeof(int);
  if (sv_alloc(c, (void **)&L.off, sizeof(int64_t) * (ns + 1)) ||
      sv_alloc(c, (void **)&L.col, I * E) || sv_alloc(c, (void **)&L.wx, D * E) ||
      sv_alloc(c, (void **)&L.wy, D * E) || sv_alloc(c, (void **)&L.wl, D * E) ||
      sv_alloc(c, (void **)&L.wh, D * E) || sv_alloc(c, (void **)&L.cnt, I * n) ||
      sv_alloc(c, (void **)&L.cnt_small, I * n) || sv_alloc(c, (void **)&L.gsx, D * n) ||
      sv_alloc(c, (void **)&L.gsy, D * n) || sv_alloc(c, (void **)&L.lsum, D * n) ||
      sv_alloc(c, (void **)&L.hsum, D * n) || sv_alloc(c, (void **)&L.vol, D * n) ||
      sv_alloc(c, (void **)&L.s, D * n) || sv_alloc(c, (void **)&L.perm, I * n) ||
      sv_alloc(c, (void **)&L.iperm, I * c->np))
    return SV_E_NOMEM;
  SV_HIP(c, hipMemcpy(L.off, off.data(), sizeof(int64_t) * (ns + 1), hipMemcpyHostToDevice));
  SV_HIP(c, hipMemcpy(L.perm, L.hperm.data(), I * n, hipMemcpyHostToDevice));
  SV_HIP(c, hipMemcpy(L.iperm, iperm.data(), I * c->np, hipMemcpyHostToDevice));
  dim3 grid(ns, (wmax + PK_K - 1) / PK_K);
  if (c->has_mirror && sv_alloc(c, (void **)&L.lflip, (size_t)L.nnz + 1)) return SV_E_NOMEM;
  k_loop_pack<<<grid, PK_T, 0, c->stream>>>(n, L.off, L.perm, L.iperm, c->row_ptr, c->cnt,
                                               c->col_raw, c->irel, c->vrel, c->wx, c->wy,
                                               c->wl, c->wh, L.nnz, L.col, L.wx, L.wy, L.wl,
                                               L.wh, L.lflip);
  SV_HIP(c, hipGetLastError());
  const int B = 256, nb = (n + B - 1) / B;
  k_permute<int><<<nb, B, 0, c->stream>>>(n, L.perm, c->cnt, L.cnt);
  k_permute<int><<<nb, B, 0, c->stream>>>(n, L.perm, c->cnt_small, L.cnt_small);
  const double *src[6] = {c->gsx, c->gsy, c->lsum, c->hsum, c->vol, c->s};
  double *dst[6] = {L.gsx, L.gsy, L.lsum, L.hsum, L.vol, L.s};
  for (int k = 0; k < 6; k++) k_permute<double><<<nb, B, 0, c->stream>>>(n, L.perm, src[k], dst[k]);
  SV_HIP(c, hipGetLastError());
  SV_HIP(c, hipStreamSynchronize(c->stream));
  if (int rc = build_halos(c)) return rc;
  L.valid = true;
  return 0;
}

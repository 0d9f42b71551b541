// k_stencil.hip -- the derivative operators of derivatives.F90 on the SELL-64
// stencil layout: one lane per node, links of a 64-node slice interleaved so a
// wavefront reads 64 consecutive indices / weights per link (coalesced), and
// the sum runs over k in the reference's order:
//   calc_gradient      :26-71    d_i = sum_k phi_j w_k - phi_i sum_k w_k
//   calc_laplacian     :73-116
//   calc_gradient_only :416-464
//   calc_filtered_var  :360-414  phi_i + (sum_k phi_j w_k - phi_i sum_k w_k)
//   calc_filter_term   :466-511
// Gradients and Laplacians run over the small stencil (ij_count_small), the
// filter over the full one (ij_count).
#include "sv_internal.h"

namespace {

struct ApplyArgs {
  int npfb;
  const int64_t *off;
  const int *kcnt;          // links per node for this operator
  const int *col;
  const double *w0, *w1;
  const double *s0, *s1;    // weight sums
  const double *phi;
  double *o0, *o1;
};

template <int OP>
__global__ __launch_bounds__(256) void k_apply(ApplyArgs a) {
  const int i = sv_xcd_block() * blockDim.x + threadIdx.x;
  if (i >= a.npfb) return;
  const int64_t b = a.off[i >> 6] + (i & 63);
  const int K = a.kcnt[i];
  double t0 = 0.0, t1 = 0.0;
  for (int k = 0; k < K; k++) {
    const int64_t e = b + (int64_t)k * SV_WAVE;
    const double pj = a.phi[a.col[e]];
    t0 = t0 + pj * a.w0[e];
    if (OP == SV_OP_GRAD) t1 = t1 + pj * a.w1[e];
  }
  const double pi = a.phi[i];
  if (OP == SV_OP_GRAD) {
    a.o0[i] = t0 - pi * a.s0[i];
    a.o1[i] = t1 - pi * a.s1[i];
  } else if (OP == SV_OP_FILTER) {
    a.o0[i] = pi + (t0 - pi * a.s0[i]);
  } else {
    a.o0[i] = t0 - pi * a.s0[i];
  }
}

}  // namespace

namespace {

// The mass correction of filter_variables (rhs_lc.F90:871-890, -Dmcorr):
// ro -= sum_i (ro_i - rho_char) vol_i / sum_i vol_i over the real nodes, with
// rho_char = 1 (load_data.F90:132).  One block: each thread sums a strided
// slice in node order, then a fixed-shape tree, so the result does not depend
// on scheduling.
__global__ __launch_bounds__(1024) void k_mass_correct(int n, const double *vol, double *ro) {
  __shared__ double sm[1024], sv[1024];
  const int t = threadIdx.x;
  double tm = 0.0, tv = 0.0;
  for (int i = t; i < n; i += 1024) {
    tm += (ro[i] - 1.0) * vol[i];
    tv += vol[i];
  }
  sm[t] = tm;
  sv[t] = tv;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (t < o) {
      sm[t] += sm[t + o];
      sv[t] += sv[t + o];
    }
    __syncthreads();
  }
  const double dro = sm[0] / sv[0];
  for (int i = t; i < n; i += 1024) ro[i] = ro[i] - dro;
}

}  // namespace

int sv_dev_mass_correct(sv_ctx *c, double *ro, hipStream_t st) {
  k_mass_correct<<<1, 1024, 0, st>>>(c->npfb, c->vol, ro);
  SV_HIP(c, hipGetLastError());
  return 0;
}

int sv_dev_apply(sv_ctx *c, int op, const double *phi, double *out, bool raw_cols,
                 hipStream_t st) {
  ApplyArgs a{};
  a.npfb = c->npfb;
  a.off = c->sell_off;
  a.col = raw_cols ? c->scol_raw : c->scol;
  a.phi = phi;
  a.o0 = out;
  a.o1 = out + c->npfb;
  switch (op) {
  case SV_OP_GRAD: a.kcnt = c->cnt_small; a.w0 = c->swx; a.w1 = c->swy; a.s0 = c->gsx; a.s1 = c->gsy; break;
  case SV_OP_LAP:
  case SV_OP_LAP_TRANSVERSE: a.kcnt = c->cnt_small; a.w0 = c->swl; a.s0 = c->lsum; break;
  case SV_OP_GRAD_X: a.kcnt = c->cnt_small; a.w0 = c->swx; a.s0 = c->gsx; break;
  case SV_OP_GRAD_Y: a.kcnt = c->cnt_small; a.w0 = c->swy; a.s0 = c->gsy; break;
  case SV_OP_FILTER:
  case SV_OP_FILTER_TERM: a.kcnt = c->cnt; a.w0 = c->swh; a.s0 = c->hsum; break;
  default: return sv_fail(c, SV_E_ARG, "unknown operator %d", op);
  }
  const int B = 256, nb = (c->npfb + B - 1) / B;
  switch (op) {
  case SV_OP_GRAD: k_apply<SV_OP_GRAD><<<nb, B, 0, st>>>(a); break;
  case SV_OP_FILTER: k_apply<SV_OP_FILTER><<<nb, B, 0, st>>>(a); break;
  default: k_apply<SV_OP_LAP><<<nb, B, 0, st>>>(a); break;
  }
  SV_HIP(c, hipGetLastError());
  // calc_laplacian_transverse_only_on_bound: the Laplacian on the internal
  // nodes (derivatives.F90:315-325), the boundary nodes overwritten with the
  // tangential part (:327-337)
  if (op == SV_OP_LAP_TRANSVERSE) return sv_dev_boundary_tt_apply(c, phi, out, raw_cols, st);
  return 0;
}

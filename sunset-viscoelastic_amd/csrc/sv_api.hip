// sv_api.hip -- the extern "C" boundary of libsunset_amd.so (declared in
// include/sunset_amd.h).  Host-side orchestration only: argument checks,
// uploads/downloads in the reference layouts, and launches of the kernels in
// k_nodes / k_weights / k_stencil / k_step.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>

#include "sv_internal.h"

// Message of the last failure that had no context to hold it (sv_create,
// sv_ensemble_create), per host thread; sv_last_error(NULL) returns it.
static thread_local std::string g_noctx_err;

int sv_fail(sv_ctx *c, int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  else g_noctx_err = buf;
  return code;
}

int sv_alloc(sv_ctx *c, void **p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return sv_fail(c, SV_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  c->allocs.push_back(*p);
  return 0;
}

void sv_free(sv_ctx *c, void *p) {
  if (!p) return;
  auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
  if (it != c->allocs.end()) c->allocs.erase(it);
  (void)hipFree(p);
}

#define SV_CHECK(c, cond, ...)                        \
  do {                                                \
    if (!(cond)) return sv_fail((c), SV_E_ARG, __VA_ARGS__); \
  } while (0)

int sv_ensure_sell(sv_ctx *c) {
  if (!c->sell_dirty) return 0;
  return sv_dev_build_sell(c);
}

int sv_ensure_loop(sv_ctx *c) {
  if (c->L.valid) return 0;
  return sv_dev_build_loop(c);
}

static int set_device(sv_ctx *c) {
  SV_HIP(c, hipSetDevice(c->device));
  return 0;
}

extern "C" {

const char *sv_version(void) { return "sunset_amd 0.1 (gfx950)"; }

static void derive_phys(sv_ctx *c);
static int upload_segment(sv_ctx *c);

int sv_create(const sv_config *cfg, sv_ctx **out) {
  if (!cfg || !out) return sv_fail(nullptr, SV_E_ARG, "sv_create: null argument");
  *out = nullptr;
  if (cfg->form != SV_FORM_DI && cfg->form != SV_FORM_LC)
    return sv_fail(nullptr, SV_E_ARG, "sv_create: form %d is neither DI (0) nor LC (1)",
                   cfg->form);
  // the orders the weight kernels are built for (common_parameter.F90:45-66
  // also allows 12)
  for (int m : {cfg->morder, cfg->forder})
    if (!sv_order_supported(m))
      return sv_fail(nullptr, SV_E_ARG,
                     "sv_create: LABFM order %d not supported on the device (4, 6, 8, 10)", m);
  sv_ctx *c = new sv_ctx();
  c->cfg = *cfg;
  c->device = cfg->device;
  hipError_t e = hipSetDevice(c->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return sv_fail(nullptr, SV_E_HIP, "sv_create: device %d: %s", cfg->device,
                   hipGetErrorString(e));
  }
  // this device's address of the allocation-free diagnostics' scratch array
  sv_diag_scratch_init(c->device);
  *out = c;
  return 0;
}

void sv_destroy(sv_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  for (void *p : c->allocs) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  sv_adj_free(c);
  delete c;
}

const char *sv_last_error(const sv_ctx *c) { return c ? c->err.c_str() : g_noctx_err.c_str(); }

int sv_sync(sv_ctx *c) {
  SV_CHECK(c, c, "null context");
  SV_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

int sv_set_domain(sv_ctx *c, double xmin, double xmax, double ymin, double ymax,
                  const int32_t bc[4], double smax, double L_char) {
  SV_CHECK(c, c && bc, "null argument");
  SV_CHECK(c, !c->is_ensemble, "not available on an ensemble context");
  SV_CHECK(c, xmax > xmin && ymax > ymin && smax > 0 && L_char > 0, "bad domain");
  const bool perx = bc[0] == 1, pery = bc[2] == 1;
  SV_CHECK(c, perx == (bc[1] == 1) && pery == (bc[3] == 1),
           "periodic boundary conditions incorrectly specified (mirror_boundaries.F90:61-69)");
  c->xmin = xmin; c->xmax = xmax; c->ymin = ymin; c->ymax = ymax;
  for (int k = 0; k < 4; k++) c->bc[k] = bc[k];
  c->smax = smax;
  // h0 = hovs*smax (setup_domain.F90:105), hovs from the filter order
  const int fo = c->cfg.forder;
  const double hovs = fo == 4 ? 1.8 : fo == 6 ? 2.4 : fo == 10 ? 3.1 : 2.7;
  c->h0 = hovs * smax;
  c->h2 = c->h0 * c->h0;
  c->L_char = L_char;
  c->have_domain = true;
  return 0;
}

int sv_set_nodes(sv_ctx *c, int32_t npfb, const double *x, const double *y, const double *s,
                 const double *h, const double *h_small, const int32_t *node_type) {
  SV_CHECK(c, c && x && y && s && h && h_small, "null argument");
  SV_CHECK(c, c->have_domain, "sv_set_domain must come first");
  SV_CHECK(c, !c->is_ensemble, "not available on an ensemble context");
  SV_CHECK(c, npfb > 0, "npfb must be positive");
  if (int rc = set_device(c)) return rc;
  for (int i = 0; i < npfb; i++)
    SV_CHECK(c, h[i] > 0 && h_small[i] > 0 && h_small[i] <= h[i] && s[i] > 0,
             "node %d: need 0 < h_small <= h and s > 0", i);
  // drop everything derived from an earlier node set
  for (void *p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
  sv_adj_free(c);
  c->x = c->y = c->h = c->hs = c->s = nullptr;
  c->node_type = c->irel = c->vrel = nullptr;
  c->cnt = c->cnt_small = nullptr; c->row_ptr = nullptr; c->col_raw = nullptr;
  c->wx = c->wy = c->wl = c->wh = nullptr;
  c->gsx = c->gsy = c->lsum = c->hsum = c->vol = nullptr;
  c->sell_off = nullptr; c->sell_w = nullptr; c->scol = c->scol_raw = nullptr;
  c->swx = c->swy = c->swl = c->swh = nullptr; c->nslices = 0; c->sell_nnz = 0;
  c->singular = nullptr; c->fx = nullptr; c->partial = nullptr; c->ctrl = nullptr;
  c->U = c->Uf = c->reg = c->eacc = nullptr;
  c->GA = c->GB = nullptr;
  c->segk = nullptr; c->blk_seg = nullptr; c->hseg.clear();
  c->L = sv_loopl{};
  c->tmp_phi = c->tmp_out = nullptr; c->tmp_cap = 0;
  if (c->graph_exec) { (void)hipGraphExecDestroy(c->graph_exec); c->graph_exec = nullptr; }
  c->graph_valid = false;
  c->have_lists = c->have_weights = c->have_sums = c->have_filter = false;
  c->have_state = false;
  c->blist = nullptr; c->rnx = c->rny = nullptr; c->ooR = nullptr;
  c->wtt = c->wtt_sum = nullptr; c->wtt_ld = 0;
  c->fields_buf = nullptr;
  c->have_normals = c->have_bweights = false;
  // boundary nodes and their FD rows (setup_domain.F90:179-193, :314-326)
  c->hnt.assign(npfb, 999);
  if (node_type) c->hnt.assign(node_type, node_type + npfb);
  c->hblist.clear();
  for (int i = 0; i < npfb; i++) {
    const int t = c->hnt[i];
    if (t >= 0 && t <= 2) {
      SV_CHECK(c, i + 4 < npfb, "boundary node %d: its four FD rows must follow it", i);
      for (int j = 1; j <= 4; j++)
        SV_CHECK(c, c->hnt[i + j] == -j,
                 "boundary node %d: node %d must be its FD row %d (node type %d)", i, i + j, j,
                 -j);
      c->hblist.push_back(i);
    } else if (t < 0) {
      SV_CHECK(c, t >= -4 && i + t >= 0 && c->hnt[i + t] >= 0 && c->hnt[i + t] <= 2,
               "node %d: FD row type %d without its boundary node", i, t);
    } else {
      SV_CHECK(c, t == 999 || t == 998, "node %d: unknown node type %d", i, t);
    }
  }
  c->nb = (int)c->hblist.size();
  c->npfb = npfb;
  c->cap = 9 * npfb + 64;      // at most 8 ghosts per node
  const size_t D = sizeof(double) * c->cap, I = sizeof(int) * c->cap;
  if (sv_alloc(c, (void **)&c->x, D) || sv_alloc(c, (void **)&c->y, D) ||
      sv_alloc(c, (void **)&c->h, D) || sv_alloc(c, (void **)&c->hs, D) ||
      sv_alloc(c, (void **)&c->s, D) || sv_alloc(c, (void **)&c->node_type, I) ||
      sv_alloc(c, (void **)&c->irel, I) || sv_alloc(c, (void **)&c->vrel, I))
    return SV_E_NOMEM;
  const size_t n8 = sizeof(double) * npfb;
  SV_HIP(c, hipMemcpy(c->x, x, n8, hipMemcpyHostToDevice));
  SV_HIP(c, hipMemcpy(c->y, y, n8, hipMemcpyHostToDevice));
  SV_HIP(c, hipMemcpy(c->s, s, n8, hipMemcpyHostToDevice));
  SV_HIP(c, hipMemcpy(c->h, h, n8, hipMemcpyHostToDevice));
  SV_HIP(c, hipMemcpy(c->hs, h_small, n8, hipMemcpyHostToDevice));
  if (node_type) {
    SV_HIP(c, hipMemcpy(c->node_type, node_type, sizeof(int) * npfb, hipMemcpyHostToDevice));
  } else {
    std::vector<int> nt(npfb, 999);
    SV_HIP(c, hipMemcpy(c->node_type, nt.data(), sizeof(int) * npfb, hipMemcpyHostToDevice));
  }
  double hm = 0, hn = 0;
  for (int i = 0; i < npfb; i++) {
    hm = h[i] > hm ? h[i] : hm;
    hn = (i == 0 || h[i] < hn) ? h[i] : hn;
  }
  c->hmax = hm;
  c->hmin = hn;
  if (c->nb > 0) {
    if (sv_alloc(c, (void **)&c->blist, sizeof(int) * c->nb)) return SV_E_NOMEM;
    SV_HIP(c, hipMemcpy(c->blist, c->hblist.data(), sizeof(int) * c->nb,
                        hipMemcpyHostToDevice));
  }
  return sv_dev_create_ghosts(c);
}

int sv_set_boundaries(sv_ctx *c, const double *rnx, const double *rny) {
  SV_CHECK(c, c && rnx && rny, "null argument");
  SV_CHECK(c, c->npfb > 0, "sv_set_nodes must come first");
  SV_CHECK(c, !c->is_ensemble, "not available on an ensemble context");
  if (int rc = set_device(c)) return rc;
  return sv_dev_boundary_normals(c, rnx, rny);
}

int sv_set_inflow(sv_ctx *c, int32_t inflow_type, int32_t velocity_control, double u_start,
                  double u_end, double ramp_time) {
  SV_CHECK(c, c, "null argument");
  SV_CHECK(c, !c->is_ensemble, "not available on an ensemble context");
  SV_CHECK(c, inflow_type >= 0 && inflow_type <= 2, "inflow type %d: 0 soft, 1 hard, 2 "
           "inflow-outflow", inflow_type);
  SV_CHECK(c, velocity_control == 0 || velocity_control == 1,
           "inflow velocity control %d: 0 fixed, 1 ramp", velocity_control);
  SV_CHECK(c, velocity_control == 0 || ramp_time > 0, "ramp_time must be positive");
  c->inflow_type = inflow_type;
  c->u_control = velocity_control;
  c->u_start = u_start;
  c->u_end = u_end;
  c->ramp_time = ramp_time;
  if (c->have_phys) {
    derive_phys(c);
    c->graph_valid = false;
    if (c->have_state) {
      if (int rc = set_device(c)) return rc;
      return upload_segment(c);
    }
  }
  return 0;
}

int sv_calc_boundary_weights(sv_ctx *c) {
  SV_CHECK(c, c && c->have_weights, "sv_calc_labf_weights must come first");
  SV_CHECK(c, !c->have_sums, "boundary weights come before sv_calc_labf_sums "
           "(sunset.F90:54-55)");
  SV_CHECK(c, c->nb == 0 || c->have_normals, "sv_set_boundaries must come first");
  SV_CHECK(c, c->nb == 0 || (c->cfg.forder <= 8 && c->cfg.morder <= 8),
           "boundary nodes with LABFM order 10 are not supported on the device (its "
           "11-point FD stencils, common_parameter.F90:111-112)");
  if (int rc = set_device(c)) return rc;
  return sv_dev_boundary_weights(c);
}

int sv_get_ghosts(sv_ctx *c, int32_t *np, double *x, double *y, int32_t *irelation,
                  int32_t *vrelation, int32_t cap) {
  SV_CHECK(c, c && c->np > 0, "no nodes");
  if (np) *np = c->np;
  const int ng = c->np - c->npfb;
  if (x || y) SV_CHECK(c, cap >= c->np, "capacity %d < np %d", cap, c->np);
  if (x) SV_HIP(c, hipMemcpy(x, c->x, sizeof(double) * c->np, hipMemcpyDeviceToHost));
  if (y) SV_HIP(c, hipMemcpy(y, c->y, sizeof(double) * c->np, hipMemcpyDeviceToHost));
  if (irelation || vrelation) {
    std::vector<int> ir(c->np), vr(c->np);
    SV_HIP(c, hipMemcpy(ir.data(), c->irel, sizeof(int) * c->np, hipMemcpyDeviceToHost));
    SV_HIP(c, hipMemcpy(vr.data(), c->vrel, sizeof(int) * c->np, hipMemcpyDeviceToHost));
    for (int g = 0; g < ng; g++) {
      if (irelation) irelation[g] = ir[c->npfb + g] + 1;
      if (vrelation) vrelation[g] = vr[c->npfb + g];
    }
  }
  return 0;
}

int sv_find_neighbours(sv_ctx *c, int32_t nplink, int32_t *ij_count, int32_t *ij_count_small,
                       int32_t *ij_link) {
  SV_CHECK(c, c && c->npfb > 0, "sv_set_nodes must come first");
  if (int rc = set_device(c)) return rc;
  if (int rc = sv_dev_find_neighbours(c)) return rc;
  c->nplink = nplink;
  if (ij_count)
    SV_HIP(c, hipMemcpy(ij_count, c->cnt, sizeof(int) * c->npfb, hipMemcpyDeviceToHost));
  if (ij_count_small)
    SV_HIP(c, hipMemcpy(ij_count_small, c->cnt_small, sizeof(int) * c->npfb,
                        hipMemcpyDeviceToHost));
  if (ij_link) {
    if (nplink < c->maxK)
      return sv_fail(c, SV_E_OVERFLOW, "a node has %d neighbours > nplink %d", c->maxK, nplink);
    std::vector<int> col(c->nnz), cnt(c->npfb);
    std::vector<int64_t> rp(c->npfb + 1);
    SV_HIP(c, hipMemcpy(col.data(), c->col_raw, sizeof(int) * c->nnz, hipMemcpyDeviceToHost));
    SV_HIP(c, hipMemcpy(rp.data(), c->row_ptr, sizeof(int64_t) * (c->npfb + 1),
                        hipMemcpyDeviceToHost));
    for (int i = 0; i < c->npfb; i++) {
      int32_t *dst = ij_link + (size_t)i * nplink;
      const int K = (int)(rp[i + 1] - rp[i]);
      for (int k = 0; k < K; k++) dst[k] = col[rp[i] + k] + 1;
      for (int k = K; k < nplink; k++) dst[k] = 0;
    }
  }
  return 0;
}

int sv_adapt_stencils(sv_ctx *c, double *h, double *h_small, int32_t *reductions) {
  SV_CHECK(c, c && c->have_lists, "sv_find_neighbours must come first");
  SV_CHECK(c, !c->have_state, "stencils are adapted before sv_set_initial (sunset.F90:40-62)");
  if (int rc = set_device(c)) return rc;
  if (int rc = sv_dev_adapt(c, reductions)) return rc;
  // the lists are rebuilt from the new h and h_small, as after
  // refine_and_finalise_domain (sunset.F90:45-52); everything built on the
  // old lists is stale
  c->have_lists = c->have_weights = c->have_sums = c->have_filter = false;
  c->have_bweights = false;
  if (int rc = sv_dev_find_neighbours(c)) return rc;
  if (h) SV_HIP(c, hipMemcpy(h, c->h, sizeof(double) * c->npfb, hipMemcpyDeviceToHost));
  if (h_small)
    SV_HIP(c, hipMemcpy(h_small, c->hs, sizeof(double) * c->npfb, hipMemcpyDeviceToHost));
  return 0;
}

int sv_calc_labf_weights(sv_ctx *c, const int32_t *order_map) {
  SV_CHECK(c, c && c->have_lists, "sv_find_neighbours must come first");
  if (int rc = set_device(c)) return rc;
  c->have_bweights = false;
  if (int rc = sv_dev_weights(c, order_map)) return rc;
  if (c->n_singular > 0)
    return sv_fail(c, SV_E_SINGULAR, "%d moment systems had a zero pivot", c->n_singular);
  return 0;
}

int sv_calc_labf_sums(sv_ctx *c) {
  SV_CHECK(c, c && c->have_weights, "sv_calc_labf_weights must come first");
  SV_CHECK(c, c->nb == 0 || c->have_bweights,
           "%d boundary nodes: sv_calc_boundary_weights must come first", c->nb);
  if (int rc = set_device(c)) return rc;
  return sv_dev_sums(c);
}

int sv_filter_coefficients(sv_ctx *c) {
  SV_CHECK(c, c && c->have_sums, "sv_calc_labf_sums must come first");
  if (int rc = set_device(c)) return rc;
  return sv_dev_filter_coeffs(c);
}

static int ensure_sell(sv_ctx *c) { return sv_ensure_sell(c); }

int sv_get_weights(sv_ctx *c, int32_t nplink, double *w_grad, double *w_lap, double *w_hyp,
                   double *gsum, double *lsum, double *hsum, double *vol) {
  SV_CHECK(c, c && c->have_weights, "no weights");
  if (int rc = set_device(c)) return rc;
  const int n = c->npfb;
  std::vector<int64_t> rp(n + 1);
  SV_HIP(c, hipMemcpy(rp.data(), c->row_ptr, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
  if (w_grad || w_lap || w_hyp) SV_CHECK(c, nplink >= c->maxK, "nplink %d < %d", nplink, c->maxK);
  auto fetch = [&](const double *src, std::vector<double> &dst) -> int {
    dst.resize(c->nnz);
    SV_HIP(c, hipMemcpy(dst.data(), src, sizeof(double) * c->nnz, hipMemcpyDeviceToHost));
    return 0;
  };
  std::vector<double> a, b;
  if (w_grad) {
    if (fetch(c->wx, a) || fetch(c->wy, b)) return SV_E_HIP;
    for (int i = 0; i < n; i++)
      for (int k = 0; k < nplink; k++) {
        const bool in = k < rp[i + 1] - rp[i];
        w_grad[((size_t)i * nplink + k) * 2 + 0] = in ? a[rp[i] + k] : 0.0;
        w_grad[((size_t)i * nplink + k) * 2 + 1] = in ? b[rp[i] + k] : 0.0;
      }
  }
  if (w_lap) {
    if (fetch(c->wl, a)) return SV_E_HIP;
    for (int i = 0; i < n; i++)
      for (int k = 0; k < nplink; k++)
        w_lap[(size_t)i * nplink + k] = k < rp[i + 1] - rp[i] ? a[rp[i] + k] : 0.0;
  }
  if (w_hyp) {
    if (fetch(c->wh, a)) return SV_E_HIP;
    for (int i = 0; i < n; i++)
      for (int k = 0; k < nplink; k++)
        w_hyp[(size_t)i * nplink + k] = k < rp[i + 1] - rp[i] ? a[rp[i] + k] : 0.0;
  }
  if (gsum) {
    SV_CHECK(c, c->have_sums, "sv_calc_labf_sums not done");
    std::vector<double> gx(n), gy(n);
    SV_HIP(c, hipMemcpy(gx.data(), c->gsx, sizeof(double) * n, hipMemcpyDeviceToHost));
    SV_HIP(c, hipMemcpy(gy.data(), c->gsy, sizeof(double) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) { gsum[2 * i] = gx[i]; gsum[2 * i + 1] = gy[i]; }
  }
  if (lsum) SV_HIP(c, hipMemcpy(lsum, c->lsum, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (hsum) SV_HIP(c, hipMemcpy(hsum, c->hsum, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (vol) SV_HIP(c, hipMemcpy(vol, c->vol, sizeof(double) * n, hipMemcpyDeviceToHost));
  return 0;
}

int sv_get_timing(sv_ctx *c, const char *what, double *ms) {
  SV_CHECK(c, c && what && ms, "null argument");
  if (!strcmp(what, "neighbours")) *ms = c->t_neigh;
  else if (!strcmp(what, "weights")) *ms = c->t_weights;
  else if (!strcmp(what, "weights_tune")) *ms = c->t_weights_tune;
  else if (!strcmp(what, "step")) *ms = c->t_step;
  else if (!strcmp(what, "systems")) *ms = c->n_systems;
  else if (!strncmp(what, "weights_variant_m", 17) && (what[17] == '4' || what[17] == '6' ||
                                                        what[17] == '8'))
    *ms = c->wvariant[(what[17] - '0') / 2 - 2];
  else if (!strcmp(what, "weights_mixed")) *ms = c->wmixed_used;
  else if (!strcmp(what, "ensemble_groups")) *ms = c->lgroup_mode;
  else if (!strcmp(what, "ensemble_groups_tune_0")) *ms = c->lgroup_ms[0];
  else if (!strcmp(what, "ensemble_groups_tune_1")) *ms = c->lgroup_ms[1];
  else if (!strncmp(what, "weights_mixed_tune_", 19) && strlen(what) == 20 && what[19] >= '0' &&
           what[19] <= '2')
    *ms = c->wmixed_ms[what[19] - '0'];
  else if (!strncmp(what, "weights_tune_m", 14) && strlen(what) == 17 && what[15] == '_' &&
           (what[14] == '4' || what[14] == '6' || what[14] == '8') && what[16] >= '0' &&
           what[16] <= '3')
    *ms = c->wtune_ms[(what[14] - '0') / 2 - 2][what[16] - '0'];
  else return sv_fail(c, SV_E_ARG, "unknown timing '%s'", what);
  return 0;
}

int sv_set_option(sv_ctx *c, const char *key, double value) {
  SV_CHECK(c, c && key, "null argument");
  if (!strcmp(key, "ensemble_groups")) {
    SV_CHECK(c, c->is_ensemble, "ensemble_groups: an ensemble context option");
    SV_CHECK(c, value == -1.0 || value == 0.0 || value == 1.0,
             "ensemble_groups: -1 (retune), 0 (one launch) or 1 (one launch per group)");
    if (int rc = set_device(c)) return rc;
    c->graph_valid = false;
    if (value < 0.0) return sv_dev_tune_groups(c, 5);
    c->lgroup_mode = (value > 0.0 && c->lgroups.size() >= 2) ? 1 : 0;
    return 0;
  }
  if (!strcmp(key, "weights_persistent")) {
    c->weights_fixed = value != 0.0 ? 1 : 0;
    c->weights_autotune = false;   // an explicit choice turns the autotuning off
  } else if (!strcmp(key, "weights_variant")) {
    SV_CHECK(c, value >= 0.0 && value <= 3.0 && value == (int)value,
             "weights_variant must be 0, 1, 2 or 3");
    c->weights_fixed = (int)value;
    c->weights_autotune = false;
  } else if (!strcmp(key, "weights_mixed")) {
    SV_CHECK(c, value == 0.0 || value == 1.0 || value == 2.0, "weights_mixed must be 0, 1 or 2");
    c->weights_mixed = (int)value;
    c->weights_autotune = false;
  } else if (!strcmp(key, "weights_autotune")) {
    c->weights_autotune = value != 0.0;
    for (int k = 0; k < 3; k++) c->wtune[k] = -1;
    c->wmixed = -1;
  }
  else if (!strcmp(key, "adj_lanes")) {
    SV_CHECK(c, value == 0.0 || value == 1.0 || value == 4.0,
             "adj_lanes must be 0 (by node count), 1 or 4");
    c->adj_lanes = (int)value;
  } else if (!strcmp(key, "halo_sweeps")) {
    SV_CHECK(c, !c->have_state, "halo_sweeps must be set before sv_set_initial");
    c->halo_sweeps = value != 0.0 ? 1 : 0;
    c->L.valid = false;
  } else if (!strcmp(key, "loop_order")) {
    SV_CHECK(c, !c->have_state, "loop_order must be set before sv_set_initial");
    c->loop_order = value != 0.0 ? 1 : 0;
    c->L.valid = false;
  }
  else return sv_fail(c, SV_E_ARG, "unknown option '%s'", key);
  return 0;
}

static int ensure_tmp(sv_ctx *c) {
  const int need = c->np > 2 * c->npfb ? c->np : 2 * c->npfb;
  if (c->tmp_cap >= need) return 0;
  if (c->tmp_phi) sv_free(c, c->tmp_phi);
  if (c->tmp_out) sv_free(c, c->tmp_out);
  if (sv_alloc(c, (void **)&c->tmp_phi, sizeof(double) * need) ||
      sv_alloc(c, (void **)&c->tmp_out, sizeof(double) * need))
    return SV_E_NOMEM;
  c->tmp_cap = need;
  return 0;
}

int sv_apply(sv_ctx *c, int32_t op, double *phi, double *out) {
  SV_CHECK(c, c && phi, "null argument");
  SV_CHECK(c, c->have_weights && c->have_sums, "weights and sums must be computed first");
  SV_CHECK(c, op >= SV_OP_GRAD && op <= SV_OP_LAP_TRANSVERSE, "bad operator");
  if ((op == SV_OP_FILTER || op == SV_OP_FILTER_TERM))
    SV_CHECK(c, c->have_filter, "sv_filter_coefficients must come first");
  if (op != SV_OP_FILTER) SV_CHECK(c, out, "null output");
  if (int rc = set_device(c)) return rc;
  if (int rc = ensure_sell(c)) return rc;
  if (int rc = ensure_tmp(c)) return rc;
  SV_HIP(c, hipMemcpy(c->tmp_phi, phi, sizeof(double) * c->np, hipMemcpyHostToDevice));
  if (int rc = sv_dev_apply(c, op, c->tmp_phi, c->tmp_out, true, c->stream)) return rc;
  SV_HIP(c, hipStreamSynchronize(c->stream));
  const size_t nout = (op == SV_OP_GRAD ? 2 : 1) * (size_t)c->npfb;
  SV_HIP(c, hipMemcpy(op == SV_OP_FILTER ? phi : out, c->tmp_out, sizeof(double) * nout,
                      hipMemcpyDeviceToHost));
  return 0;
}

int sv_calc_gradient(sv_ctx *c, const double *phi, double *g) {
  return sv_apply(c, SV_OP_GRAD, const_cast<double *>(phi), g);
}
int sv_calc_laplacian(sv_ctx *c, const double *phi, double *l) {
  return sv_apply(c, SV_OP_LAP, const_cast<double *>(phi), l);
}
int sv_calc_gradient_only(sv_ctx *c, int32_t dim, const double *phi, double *g) {
  SV_CHECK(c, c, "null context");
  if (dim == 3) {  // 2D: third component is zero (derivatives.F90:455-458)
    SV_CHECK(c, g, "null output");
    memset(g, 0, sizeof(double) * c->npfb);
    return 0;
  }
  SV_CHECK(c, dim == 1 || dim == 2, "dim must be 1, 2 or 3");
  return sv_apply(c, dim == 1 ? SV_OP_GRAD_X : SV_OP_GRAD_Y, const_cast<double *>(phi), g);
}
int sv_calc_filtered_var(sv_ctx *c, double *phi) { return sv_apply(c, SV_OP_FILTER, phi, nullptr); }
int sv_calc_laplacian_transverse_only_on_bound(sv_ctx *c, const double *phi, double *l) {
  return sv_apply(c, SV_OP_LAP_TRANSVERSE, const_cast<double *>(phi), l);
}
int sv_calc_filter_term(sv_ctx *c, const double *phi, double *f) {
  return sv_apply(c, SV_OP_FILTER_TERM, const_cast<double *>(phi), f);
}

int sv_apply_dev(sv_ctx *c, int32_t op, int32_t nfields, const double *const *phi,
                 double *const *out, void *stream) {
  SV_CHECK(c, c && phi && out && nfields >= 0, "null argument");
  SV_CHECK(c, c->have_weights && c->have_sums, "weights and sums must be computed first");
  if (int rc = set_device(c)) return rc;
  if (int rc = ensure_sell(c)) return rc;
  // queued on the caller's stream as given: NULL is the null (default)
  // stream, which orders the launches with the caller's work there (a
  // framework's default stream is the null stream)
  hipStream_t st = (hipStream_t)stream;
  for (int f = 0; f < nfields; f++)
    if (int rc = sv_dev_apply(c, op, phi[f], out[f], false, st)) return rc;
  return 0;
}

static void derive_phys(sv_ctx *c) {
  const sv_phys &p = c->phys;
  sv_physk &k = c->pk;
  k.rho_char = 1.0;                                    // load_data.F90:132-148
  k.visc_total = k.rho_char * p.U_char * p.L_char / p.Re;
  k.coef_solvent = p.beta * k.visc_total;
  const double visc_polymeric = (1.0 - p.beta) * k.visc_total;
  k.lambda = p.Wi * p.L_char / p.U_char;
  k.coef_polymeric = visc_polymeric / k.lambda;
  k.csq = (p.U_char / p.Ma) * (p.U_char / p.Ma);
  k.cs = sqrt(k.csq);
  k.eps = p.epsPTT;
  k.Mdiff = p.Mdiff;
  k.bfx = p.grav[0] + p.driving_force[0];
  k.bfy = p.grav[1] + p.driving_force[1];
  k.kol_amp = p.kol_amp;
  k.kol_k = p.kol_k;
  k.U_char = p.U_char;
  k.L_char = p.L_char;
  k.err_norm_rou = 1.0 / (1.0 * p.U_char);
  k.inflow_type = c->inflow_type;
  k.u_control = c->u_control;
  k.u_start = c->u_start;
  k.u_end = c->u_end;
  k.ramp_time = c->ramp_time;
  k.Lx = (c->xmax - c->xmin) * p.L_char;
  k.p_ref = k.rho_char * k.csq;
}

static int ensure_state_buffers(sv_ctx *c) {
  if (c->U) return 0;
  c->gs = c->cfg.form == SV_FORM_LC ? SV_GS_LC : SV_GS_DI;
  c->nseg = 1;
  c->nblk = (c->npfb + SV_BLOCK - 1) / SV_BLOCK;
  c->ntot = c->nblk * SV_BLOCK;
  const size_t nu = sizeof(double) * SV_US * (size_t)c->ntot;
  const size_t ng = sizeof(double) * c->gs * (size_t)c->ntot;
  if (sv_alloc(c, (void **)&c->U, nu) || sv_alloc(c, (void **)&c->Uf, nu) ||
      sv_alloc(c, (void **)&c->reg, nu) || sv_alloc(c, (void **)&c->eacc, nu) ||
      sv_alloc(c, (void **)&c->GA, ng) || sv_alloc(c, (void **)&c->GB, ng))
    return SV_E_NOMEM;
  SV_HIP(c, hipMemset(c->GA, 0, ng));
  SV_HIP(c, hipMemset(c->GB, 0, ng));
  c->npartial = 4 * c->nblk;
  if (sv_alloc(c, (void **)&c->partial, sizeof(double) * (c->npartial + 8)) ||
      sv_alloc(c, (void **)&c->ctrl, sizeof(sv_ctrl)) ||
      sv_alloc(c, (void **)&c->fx, sizeof(double) * c->ntot) ||
      sv_alloc(c, (void **)&c->segk, sizeof(sv_segk)) ||
      sv_alloc(c, (void **)&c->blk_seg, sizeof(int) * c->nblk))
    return SV_E_NOMEM;
  SV_HIP(c, hipMemset(c->fx, 0, sizeof(double) * c->ntot));
  SV_HIP(c, hipMemset(c->blk_seg, 0, sizeof(int) * c->nblk));
  return 0;
}

// The single segment of a plain context: physics constants, volume sum and the
// per-node Kolmogorov forcing (setup only).
static int upload_segment(sv_ctx *c) {
  const int n = c->npfb;
  std::vector<double> y(n), vol(n), fx(n);
  SV_HIP(c, hipMemcpy(y.data(), c->y, sizeof(double) * n, hipMemcpyDeviceToHost));
  SV_HIP(c, hipMemcpy(vol.data(), c->vol, sizeof(double) * n, hipMemcpyDeviceToHost));
  double tv = 0.0;
  for (int i = 0; i < n; i++) tv += vol[i];
  c->volsum = tv;
  const std::vector<int> &P = c->L.hperm;   // loop order
  for (int k = 0; k < n; k++) fx[k] = c->pk.kol_amp * sin(c->pk.kol_k * y[P[k]]);
  SV_HIP(c, hipMemcpy(c->fx, fx.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  c->hseg.assign(1, sv_segk{});
  sv_segk &S = c->hseg[0];
  S.base = 0; S.npfb = n; S.blk0 = 0; S.nblk = c->nblk; S.volsum = tv; S.pk = c->pk;
  SV_HIP(c, hipMemcpy(c->segk, &S, sizeof(S), hipMemcpyHostToDevice));
  c->any_forcing = c->pk.kol_amp != 0.0;
  c->any_mdiff = c->pk.Mdiff != 0.0;
  return 0;
}

int sv_set_phys(sv_ctx *c, const sv_phys *p) {
  SV_CHECK(c, c && p, "null argument");
  SV_CHECK(c, p->Re > 0 && p->Wi > 0 && p->Ma > 0 && p->U_char > 0 && p->L_char > 0,
           "Re, Wi, Ma, U_char and L_char must be positive");
  SV_CHECK(c, !c->is_ensemble, "ensemble context: physics is set per member before "
                                 "sv_ensemble_create");
  c->phys = *p;
  derive_phys(c);
  c->have_phys = true;
  c->graph_valid = false;
  if (c->have_state) {
    if (int rc = set_device(c)) return rc;
    return upload_segment(c);
  }
  return 0;
}

// log_conf_psi_from_c on the host (setup only), conf_transforms.F90:177-215
static void psi_from_c(double a, double b, double d, double *pxx, double *pxy, double *pyy) {
  double l1, l2, R[4];
  if (fabs(b) <= 1.0e-16) {
    l1 = a; l2 = d; R[0] = 1; R[1] = 0; R[2] = 0; R[3] = 1;
  } else {
    const double tr = a + d, det = a * d - b * b;
    double t = 0.25 * tr * tr - det;
    if (t <= 1.0e-16) t = 0.0;
    t = sqrt(t);
    l1 = 0.5 * tr + t;
    l2 = 0.5 * tr - t;
    double r11 = l1 - d, r21 = b;
    const double nrm = sqrt(r11 * r11 + r21 * r21);
    r11 /= nrm; r21 /= nrm;
    R[0] = r11; R[1] = r21; R[2] = -r21; R[3] = r11;
  }
  const double e1 = log(l1), e2 = log(l2);
  const double i11 = e1 * R[0], i12 = e1 * R[1], i21 = e2 * R[2], i22 = e2 * R[3];
  *pxx = R[0] * i11 + R[2] * i21;
  *pxy = R[1] * i11 + R[3] * i21;
  *pyy = R[1] * i12 + R[3] * i22;
}

int sv_set_initial(sv_ctx *c, const double *ro, const double *u, const double *v,
                   const double *cxx, const double *cxy, const double *cyy) {
  SV_CHECK(c, c && ro && u && v && cxx && cxy && cyy, "null argument");
  SV_CHECK(c, !c->is_ensemble, "not available on an ensemble context");
  SV_CHECK(c, c->have_phys, "sv_set_phys must come first");
  SV_CHECK(c, c->have_filter, "weights, sums and filter coefficients must come first");
  if (int rc = set_device(c)) return rc;
  if (int rc = sv_ensure_loop(c)) return rc;
  SV_CHECK(c, !c->has_mirror || c->L.halo,
           "symmetric/no-slip mirror ghosts need the LDS halo sweeps (halo_sweeps = 1 and "
           "block halos within %d bytes)", SV_HALO_LDS_MAX);
  if (int rc = ensure_state_buffers(c)) return rc;
  const int n = c->npfb;
  std::vector<double> a(n), b(n), d(n), e(n), f(n), g(n);
  for (int i = 0; i < n; i++) {
    SV_CHECK(c, ro[i] > 0, "node %d: density must be positive", i);
    a[i] = ro[i];
    b[i] = ro[i] * u[i];
    d[i] = ro[i] * v[i];
    if (c->cfg.form == SV_FORM_LC) {
      SV_CHECK(c, cxx[i] * cyy[i] - cxy[i] * cxy[i] > 0 && cxx[i] > 0,
               "node %d: conformation tensor not positive definite", i);
      psi_from_c(cxx[i], cxy[i], cyy[i], &e[i], &f[i], &g[i]);
    } else {
      e[i] = cxx[i]; f[i] = cxy[i]; g[i] = cyy[i];
    }
  }
  std::vector<double> rec((size_t)SV_US * n, 0.0);
  const std::vector<int> &P = c->L.hperm;   // record k holds the caller's node P[k]
  for (int k = 0; k < n; k++) {
    double *r = rec.data() + (size_t)SV_US * k;
    const int i = P[k];
    r[0] = a[i]; r[1] = b[i]; r[2] = d[i]; r[3] = e[i]; r[4] = f[i]; r[5] = g[i];
  }
  SV_HIP(c, hipMemcpy(c->U, rec.data(), sizeof(double) * rec.size(), hipMemcpyHostToDevice));
  if (c->nb > 0) {
    // boundary records and the loop slot -> record map; u_inflow_local starts
    // at U_char, on inflow nodes at the initial u (setup_flow.F90:48-51, :130-137)
    std::vector<double> nx(n), ny(n), sp(n);
    SV_HIP(c, hipMemcpy(nx.data(), c->rnx, sizeof(double) * n, hipMemcpyDeviceToHost));
    SV_HIP(c, hipMemcpy(ny.data(), c->rny, sizeof(double) * n, hipMemcpyDeviceToHost));
    SV_HIP(c, hipMemcpy(sp.data(), c->s, sizeof(double) * n, hipMemcpyDeviceToHost));
    std::vector<double> ooR(c->nb, 0.0);
    if (c->ooR)
      SV_HIP(c, hipMemcpy(ooR.data(), c->ooR, sizeof(double) * c->nb, hipMemcpyDeviceToHost));
    std::vector<int> slot(n);
    for (int k = 0; k < n; k++) slot[P[k]] = k;
    std::vector<double> br((size_t)SV_BREC * c->nb, 0.0);
    std::vector<int> jbof(n, -1), bidx(n, -1);
    for (int jb = 0; jb < c->nb; jb++) {
      const int i = c->hblist[jb];
      jbof[i] = jb;
      double *r = br.data() + (size_t)SV_BREC * jb;
      r[0] = c->hnt[i];
      r[1] = nx[i];
      r[2] = ny[i];
      r[3] = c->hnt[i] == 1 ? u[i] : c->phys.U_char;
      for (int k = 1; k <= 4; k++) r[3 + k] = slot[i + k];   // FD rows follow the node
      r[8] = sp[i];
      r[9] = ooR[jb];
      r[10] = c->L_char;
    }
    for (int k = 0; k < n; k++) bidx[k] = jbof[P[k]];
    sv_free(c, c->L.bidx);
    sv_free(c, c->L.brec);
    c->L.bidx = nullptr;
    c->L.brec = nullptr;
    if (sv_alloc(c, (void **)&c->L.bidx, sizeof(int) * c->ntot) ||
        sv_alloc(c, (void **)&c->L.brec, sizeof(double) * br.size()))
      return SV_E_NOMEM;
    SV_HIP(c, hipMemset(c->L.bidx, 0xff, sizeof(int) * c->ntot));
    SV_HIP(c, hipMemcpy(c->L.bidx, bidx.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    SV_HIP(c, hipMemcpy(c->L.brec, br.data(), sizeof(double) * br.size(),
                        hipMemcpyHostToDevice));
  }
  if (int rc = upload_segment(c)) return rc;
  sv_ctrl h{};
  h.dt = 1.0e-6;                         // setup_flow.F90:149-151
  h.emax_np1 = h.emax_n = h.emax_nm1 = SV_PID_TOL;
  h.dt_mode = SV_DT_PID;
  SV_HIP(c, hipMemcpy(c->ctrl, &h, sizeof(h), hipMemcpyHostToDevice));
  if (int rc = sv_dev_init_state(c)) return rc;
  c->have_state = true;
  c->graph_valid = false;
  return 0;
}

int sv_set_state(sv_ctx *c, const double *ro, const double *rou, const double *rov,
                 const double *xx, const double *xy, const double *yy) {
  SV_CHECK(c, c && ro && rou && rov && xx && xy && yy, "null argument");
  SV_CHECK(c, !c->is_ensemble, "not available on an ensemble context");
  SV_CHECK(c, c->have_state, "sv_set_initial must come first (boundary records, inflow velocity)");
  if (int rc = set_device(c)) return rc;
  const int n = c->npfb;
  for (int i = 0; i < n; i++) {
    SV_CHECK(c, ro[i] > 0, "node %d: density must be positive", i);
    if (c->cfg.form == SV_FORM_DI)
      SV_CHECK(c, xx[i] * yy[i] - xy[i] * xy[i] > 0 && xx[i] > 0,
               "node %d: conformation tensor not positive definite", i);
  }
  // the six arrays to the device as they are, then into the loop-order
  // registers there (record k holds the caller's node perm[k])
  if (!c->fields_buf && sv_alloc(c, (void **)&c->fields_buf, sizeof(double) * 12 * (size_t)n))
    return SV_E_NOMEM;
  const double *src[6] = {ro, rou, rov, xx, xy, yy};
  const double *dev[6];
  for (int q = 0; q < 6; q++) {
    double *d = c->fields_buf + (size_t)q * n;
    SV_HIP(c, hipMemcpy(d, src[q], sizeof(double) * n, hipMemcpyHostToDevice));
    dev[q] = d;
  }
  if (int rc = sv_dev_gather_records(c, c->U, dev)) return rc;
  return sv_dev_init_state(c);
}

extern "C++" int sv_check_device_ptr(sv_ctx *c, const void *p, const char *what) {
  if (!p) return sv_fail(c, SV_E_ARG, "%s: null device pointer", what);
  hipPointerAttribute_t at{};
  const hipError_t e = hipPointerGetAttributes(&at, p);
  if (e != hipSuccess) (void)hipGetLastError();
  if (e != hipSuccess || at.type != hipMemoryTypeDevice || at.device != c->device)
    return sv_fail(c, SV_E_ARG,
                   "%s: %p is not device memory of GPU %d in this library's HIP runtime (a "
                   "Python process must import torch before loading the library for the two "
                   "to share one runtime)", what, p, c->device);
  return 0;
}

int sv_set_state_device(sv_ctx *c, const double *const U[6], void *stream) {
  SV_CHECK(c, c && U, "null argument");
  SV_CHECK(c, !c->is_ensemble, "not available on an ensemble context");
  SV_CHECK(c, c->have_state, "sv_set_initial must come first (boundary records, inflow velocity)");
  if (int rc = set_device(c)) return rc;
  for (int q = 0; q < 6; q++)
    if (int rc = sv_check_device_ptr(c, U[q], "sv_set_state_device")) return rc;
  // the caller's stream, the null stream included (a framework's default
  // stream): its pending work on the arrays completes first
  SV_HIP(c, hipStreamSynchronize((hipStream_t)stream));
  if (int rc = sv_dev_gather_records(c, c->U, U)) return rc;
  if (int rc = sv_dev_init_state(c)) return rc;   // synchronises the context's stream
  return 0;
}

int sv_get_state_device(sv_ctx *c, double *const U[6], void *stream) {
  SV_CHECK(c, c && U, "null argument");
  SV_CHECK(c, !c->is_ensemble, "not available on an ensemble context");
  SV_CHECK(c, c->have_state, "no state");
  if (int rc = set_device(c)) return rc;
  for (int q = 0; q < 6; q++)
    if (U[q])
      if (int rc = sv_check_device_ptr(c, U[q], "sv_get_state_device")) return rc;
  // the caller's stream, the null stream included (a framework's default
  // stream): its pending work on the arrays completes first
  SV_HIP(c, hipStreamSynchronize((hipStream_t)stream));
  if (int rc = sv_dev_scatter_records(c, c->U, U)) return rc;
  SV_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

int sv_filter_variables(sv_ctx *c, double *ro, double *rou, double *rov, double *xx, double *xy,
                        double *yy) {
  SV_CHECK(c, c && ro && rou && rov && xx && xy && yy, "null argument");
  SV_CHECK(c, c->have_filter, "weights, sums and filter coefficients must come first");
  if (int rc = set_device(c)) return rc;
  if (int rc = ensure_sell(c)) return rc;
  const int np = c->np, n = c->npfb;
  DevTmp<double> in, out;
  SV_HIP(c, in.alloc(6 * (size_t)np));
  SV_HIP(c, out.alloc(6 * (size_t)n));
  double *f[6] = {ro, rou, rov, xx, xy, yy};
  for (int k = 0; k < 6; k++) {
    SV_HIP(c, hipMemcpyAsync(in.p + (size_t)k * np, f[k], sizeof(double) * np,
                             hipMemcpyHostToDevice, c->stream));
    // calc_filtered_var on the caller's ghost values (raw ghost columns)
    if (int rc = sv_dev_apply(c, SV_OP_FILTER, in.p + (size_t)k * np, out.p + (size_t)k * n,
                              true, c->stream))
      return rc;
  }
  if (int rc = sv_dev_mass_correct(c, out.p, c->stream)) return rc;
  for (int k = 0; k < 6; k++)
    SV_HIP(c, hipMemcpyAsync(f[k], out.p + (size_t)k * n, sizeof(double) * n,
                             hipMemcpyDeviceToHost, c->stream));
  SV_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

int sv_set_control(sv_ctx *c, double dt, double time, int32_t itime, double e1, double e2,
                   double e3, int32_t dt_mode) {
  SV_CHECK(c, c && c->have_state, "sv_set_initial must come first");
  SV_CHECK(c, !c->is_ensemble, "not available on an ensemble context");
  SV_CHECK(c, dt > 0 && e1 > 0 && e2 > 0 && e3 > 0, "dt and error norms must be positive");
  SV_CHECK(c, dt_mode == SV_DT_PID || dt_mode == SV_DT_FIXED || dt_mode == SV_DT_CFL,
           "bad dt mode %d (0 PID, 1 fixed, 2 CFL-limited)", dt_mode);
  if (int rc = set_device(c)) return rc;
  sv_ctrl h{};
  SV_HIP(c, hipMemcpy(&h, c->ctrl, sizeof(h), hipMemcpyDeviceToHost));
  h.dt = dt; h.time = time; h.itime = itime;
  h.emax_np1 = e1; h.emax_n = e2; h.emax_nm1 = e3;
  h.dt_mode = dt_mode;
  SV_HIP(c, hipMemcpy(c->ctrl, &h, sizeof(h), hipMemcpyHostToDevice));
  return 0;
}

int sv_ensemble_get_control(sv_ctx *c, int32_t member, double *dt, double *time,
                            int32_t *itime, double *emax3, double *dt_cfl, double *dt_par);

int sv_get_control(sv_ctx *c, double *dt, double *time, int32_t *itime, double *emax3,
                   double *dt_cfl, double *dt_par) {
  return sv_ensemble_get_control(c, 0, dt, time, itime, emax3, dt_cfl, dt_par);
}

int sv_ensemble_get_control(sv_ctx *c, int32_t member, double *dt, double *time,
                            int32_t *itime, double *emax3, double *dt_cfl, double *dt_par) {
  SV_CHECK(c, c && c->have_state, "no state");
  SV_CHECK(c, member >= 0 && member < c->nseg, "member %d out of range [0, %d)", member,
           c->nseg);
  if (int rc = set_device(c)) return rc;
  SV_HIP(c, hipStreamSynchronize(c->stream));
  sv_ctrl h{};
  SV_HIP(c, hipMemcpy(&h, c->ctrl + member, sizeof(h), hipMemcpyDeviceToHost));
  if (dt) *dt = h.dt;
  if (time) *time = h.time;
  if (itime) *itime = h.itime;
  if (emax3) { emax3[0] = h.emax_np1; emax3[1] = h.emax_n; emax3[2] = h.emax_nm1; }
  if (dt_cfl) *dt_cfl = h.dt_cfl;
  if (dt_par) *dt_par = h.dt_par;
  if (h.flags & 1) return sv_fail(c, SV_E_STATE, "non-finite value in the state");
  return 0;
}

int sv_calc_all_rhs(sv_ctx *c, int32_t iRKstep, double *r0, double *r1, double *r2, double *r3,
                    double *r4, double *r5) {
  SV_CHECK(c, c && c->have_state, "sv_set_initial must come first");
  SV_CHECK(c, !c->is_ensemble, "not available on an ensemble context");
  if (int rc = set_device(c)) return rc;
  const size_t n = c->npfb;
  // the right-hand sides in loop order in the first half of the staging
  // area, put in the caller's order in the second half on the device
  if (!c->fields_buf && sv_alloc(c, (void **)&c->fields_buf, sizeof(double) * 12 * n))
    return SV_E_NOMEM;
  double *lo = c->fields_buf, *co = c->fields_buf + 6 * n;
  if (int rc = sv_dev_rhs(c, iRKstep, lo)) return rc;
  if (int rc = sv_dev_to_caller_order(c, 6, lo, co)) return rc;
  SV_HIP(c, hipStreamSynchronize(c->stream));
  double *dst[6] = {r0, r1, r2, r3, r4, r5};
  for (int f = 0; f < 6; f++)
    if (dst[f])
      SV_HIP(c, hipMemcpy(dst[f], co + (size_t)f * n, sizeof(double) * n, hipMemcpyDeviceToHost));
  return 0;
}

int sv_run(sv_ctx *c, int32_t nsteps, int32_t use_graph) {
  SV_CHECK(c, c && c->have_state, "sv_set_initial must come first");
  SV_CHECK(c, nsteps >= 0, "nsteps must be >= 0");
  if (int rc = set_device(c)) return rc;
  EventTmp e0, e1;
  SV_HIP(c, e0.create());
  SV_HIP(c, e1.create());
  SV_HIP(c, hipEventRecord(e0, c->stream));
  if (use_graph && nsteps > 0) {
    if (!c->graph_valid) {
      if (c->graph_exec) { (void)hipGraphExecDestroy(c->graph_exec); c->graph_exec = nullptr; }
      hipGraph_t g;
      SV_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int rc = sv_dev_step(c, nullptr);
      hipError_t ce = hipStreamEndCapture(c->stream, &g);
      if (rc) return rc;
      SV_HIP(c, ce);
      SV_HIP(c, hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0));
      SV_HIP(c, hipGraphDestroy(g));
      c->graph_valid = true;
    }
    for (int s = 0; s < nsteps; s++) SV_HIP(c, hipGraphLaunch(c->graph_exec, c->stream));
  } else {
    for (int s = 0; s < nsteps; s++)
      if (int rc = sv_dev_step(c, nullptr)) return rc;
  }
  SV_HIP(c, hipEventRecord(e1, c->stream));
  SV_HIP(c, hipEventSynchronize(e1));
  float ms = 0;
  SV_HIP(c, hipEventElapsedTime(&ms, e0, e1));
  c->t_step = ms;
  return 0;
}

int sv_run_profiled(sv_ctx *c, int32_t nsteps, double *kernel_ms, double *total_ms) {
  SV_CHECK(c, c && c->have_state, "sv_set_initial must come first");
  SV_CHECK(c, nsteps > 0, "nsteps must be > 0");
  if (int rc = set_device(c)) return rc;
  const int NE = SV_NKERNELS + 1;
  std::vector<EventTmp> evs((size_t)nsteps * NE);
  for (auto &e : evs) SV_HIP(c, e.create());
  std::vector<hipEvent_t> ev(evs.size());
  for (size_t k = 0; k < evs.size(); k++) ev[k] = evs[k];
  for (int s = 0; s < nsteps; s++)
    if (int rc = sv_dev_step(c, ev.data() + (size_t)s * NE)) return rc;
  SV_HIP(c, hipEventSynchronize(ev.back()));
  double acc[SV_NKERNELS] = {0};
  for (int s = 0; s < nsteps; s++)
    for (int k = 0; k < SV_NKERNELS; k++) {
      float ms = 0;
      SV_HIP(c, hipEventElapsedTime(&ms, ev[(size_t)s * NE + k], ev[(size_t)s * NE + k + 1]));
      acc[k] += ms;
    }
  float tot = 0;
  SV_HIP(c, hipEventElapsedTime(&tot, ev.front(), ev.back()));
  for (int k = 0; k < SV_NKERNELS; k++)
    if (kernel_ms) kernel_ms[k] = acc[k] / nsteps;
  if (total_ms) *total_ms = tot;
  c->t_step = tot;
  return 0;
}

int sv_get_stencil_stats(sv_ctx *c, double *stats) {
  SV_CHECK(c, c && c->have_lists && stats, "no stencils");
  std::vector<int> cnt(c->npfb), cs(c->npfb);
  SV_HIP(c, hipMemcpy(cnt.data(), c->cnt, sizeof(int) * c->npfb, hipMemcpyDeviceToHost));
  SV_HIP(c, hipMemcpy(cs.data(), c->cnt_small, sizeof(int) * c->npfb, hipMemcpyDeviceToHost));
  double a = 0, b = 0;
  int mx = 0;
  for (int i = 0; i < c->npfb; i++) { a += cnt[i]; b += cs[i]; mx = cnt[i] > mx ? cnt[i] : mx; }
  stats[0] = a / c->npfb;          // mean ij_count
  stats[1] = b / c->npfb;          // mean ij_count_small
  stats[2] = mx;                   // max ij_count
  stats[3] = (double)(c->L.valid ? c->L.nnz : c->sell_nnz);  // SELL entries incl. padding
  stats[4] = c->np - c->npfb;      // ghosts
  stats[5] = 0.0;                  // halo records per node staged by the sweeps
  if (c->L.valid && c->L.halo) {
    int64_t tot = 0;
    const int nb = (c->npfb + SV_BLOCK - 1) / SV_BLOCK;
    SV_HIP(c, hipMemcpy(&tot, c->L.hoff + nb, sizeof(int64_t), hipMemcpyDeviceToHost));
    stats[5] = (double)tot / c->npfb;
  }
  return 0;
}

int sv_ensemble_get_state(sv_ctx *c, int32_t member, double *ro, double *rou, double *rov,
                          double *u, double *v, double *pxx, double *pxy, double *pyy,
                          double *cxx, double *cxy, double *cyy);

int sv_get_state(sv_ctx *c, double *ro, double *rou, double *rov, double *u, double *v,
                 double *pxx, double *pxy, double *pyy, double *cxx, double *cxy, double *cyy) {
  return sv_ensemble_get_state(c, 0, ro, rou, rov, u, v, pxx, pxy, pyy, cxx, cxy, cyy);
}

int sv_ensemble_get_state(sv_ctx *c, int32_t member, double *ro, double *rou, double *rov,
                          double *u, double *v, double *pxx, double *pxy, double *pyy,
                          double *cxx, double *cxy, double *cyy) {
  SV_CHECK(c, c && c->have_state, "no state");
  SV_CHECK(c, member >= 0 && member < c->nseg, "member %d out of range [0, %d)", member,
           c->nseg);
  if (int rc = set_device(c)) return rc;
  const sv_segk &S = c->hseg[member];
  const double *U = c->U + (size_t)S.base * SV_US;
  const double *GA = c->GA + (size_t)S.base * c->gs;
  const std::vector<int> &P = c->is_ensemble ? c->member_perm[member] : c->L.hperm;
  const size_t n = (size_t)S.npfb;
  if (!c->is_ensemble && c->L.perm) {
    // fields picked and put in the caller's order on the device, then one
    // copy per field (the host scatter of eleven fields cost ~50 ms at 1M)
    if (!c->fields_buf && sv_alloc(c, (void **)&c->fields_buf, sizeof(double) * 12 * n))
      return SV_E_NOMEM;
    if (int rc = sv_dev_get_fields(c, c->fields_buf)) return rc;
    SV_HIP(c, hipStreamSynchronize(c->stream));
    double *dst[11] = {ro, rou, rov, u, v, pxx, pxy, pyy, cxx, cxy, cyy};
    const bool lc_ = c->cfg.form == SV_FORM_LC;
    for (int q = 0; q < 11; q++)
      if (dst[q] && (lc_ || q < 5 || q > 7))
        SV_HIP(c, hipMemcpy(dst[q], c->fields_buf + (size_t)q * n, sizeof(double) * n,
                            hipMemcpyDeviceToHost));
    return 0;
  }
  SV_HIP(c, hipStreamSynchronize(c->stream));
  // two contiguous copies of the member's records, fields picked on the host
  // (a strided copy of 8-byte rows is far slower than the whole record array)
  std::vector<double> hu(n * SV_US), hg(n * c->gs);
  SV_HIP(c, hipMemcpy(hu.data(), U, sizeof(double) * hu.size(), hipMemcpyDeviceToHost));
  SV_HIP(c, hipMemcpy(hg.data(), GA, sizeof(double) * hg.size(), hipMemcpyDeviceToHost));
  const bool lc = c->cfg.form == SV_FORM_LC;
  struct { double *dst; const double *src; int stride; } m[] = {
      {ro, hu.data() + 0, SV_US}, {rou, hu.data() + 1, SV_US}, {rov, hu.data() + 2, SV_US},
      {u, hg.data() + SV_G_U, c->gs}, {v, hg.data() + SV_G_V, c->gs},
      {pxx, lc ? hu.data() + 3 : nullptr, SV_US}, {pxy, lc ? hu.data() + 4 : nullptr, SV_US},
      {pyy, lc ? hu.data() + 5 : nullptr, SV_US},
      {cxx, hg.data() + SV_G_CXX, c->gs}, {cxy, hg.data() + SV_G_CXY, c->gs},
      {cyy, hg.data() + SV_G_CYY, c->gs}};
  for (auto &e : m)
    if (e.dst && e.src)
      for (size_t k = 0; k < n; k++) e.dst[P[k]] = e.src[k * e.stride];   // loop order -> caller's
  return 0;
}

}  // extern "C"

// sv_ensemble.hip -- one device context holding several independent simulations
// (members of different sizes, orders and physics) so that one launch per
// kernel, and one captured hipGraph per step, advances all of them.
//
// Layout: every member's per-node arrays (SELL stencils, sums, state records)
// are concatenated, member m starting at node base_m, a multiple of SV_BLOCK,
// with its SELL entries offset and its column indices rebased.  Each block of
// the node kernels therefore belongs to one member; the per-member reductions
// (time step, error norm, mass correction) run one block per member over that
// member's block range, and each member keeps its own control block (dt, PID
// history, itime, inflow ramp) and physics constants.  Boundary records are
// concatenated with their FD-row slots rebased; mirror ghosts keep their
// sign-flip codes in the rebased halo keys.  The members' own contexts stay
// valid and unchanged: the ensemble holds copies taken at creation.
#include <string.h>

#include <algorithm>

#include "sv_internal.h"

namespace {

int copy_d2d(sv_ctx *E, void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return 0;
  SV_HIP(E, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, E->stream));
  return 0;
}

}  // namespace

extern "C" {

int sv_ensemble_create(sv_ctx *const *members, int32_t nmembers, sv_ctx **out) {
  // The ensemble context is not returned on failure, so every refusal is
  // reported as the context-free message (sv_last_error(NULL)) and copied to
  // members[0]'s message when there is a first member.
  sv_ctx *m0 = members && nmembers >= 1 ? members[0] : nullptr;
  auto refuse = [&](int rc, const char *msg, int m) {
    if (m >= 0) sv_fail(nullptr, rc, "sv_ensemble_create: member %d: %s", m, msg);
    else sv_fail(nullptr, rc, "sv_ensemble_create: %s", msg);
    if (m0) m0->err = sv_last_error(nullptr);
    return rc;
  };
  if (!out) return refuse(SV_E_ARG, "null output pointer", -1);
  *out = nullptr;
  if (!members || nmembers < 1) return refuse(SV_E_ARG, "no members", -1);
  if (!m0) return refuse(SV_E_ARG, "null member", 0);
  for (int m = 0; m < nmembers; m++) {
    const sv_ctx *c = members[m];
    if (!c) return refuse(SV_E_ARG, "null member", m);
    if (c->is_ensemble) return refuse(SV_E_ARG, "is an ensemble context", m);
    if (!c->have_state) return refuse(SV_E_ARG, "has no state (sv_set_initial first)", m);
    if (c->cfg.form != m0->cfg.form)
      return refuse(SV_E_ARG, "conformation form differs from member 0's", m);
    if (c->device != m0->device) return refuse(SV_E_ARG, "device differs from member 0's", m);
  }
  sv_config cfg = m0->cfg;
  sv_ctx *E = nullptr;
  if (int rc = sv_create(&cfg, &E)) return refuse(rc, sv_last_error(nullptr), -1);
  // E's own message (allocation or HIP failures) goes to the caller
  auto fail = [&](int rc) {
    std::string msg = E->err.empty() ? std::string("failed") : E->err;
    sv_destroy(E);
    return refuse(rc, msg.c_str(), -1);
  };
  auto fail_msg = [&](int rc, const char *msg) {
    E->err = msg;
    return fail(rc);
  };
  E->is_ensemble = true;
  E->nseg = nmembers;
  E->gs = m0->gs;
  // layout
  E->hseg.resize(nmembers);
  std::vector<int64_t> sell_base(nmembers);
  int64_t nodes = 0, sell = 0;
  int blocks = 0;
  for (int m = 0; m < nmembers; m++) {
    sv_ctx *c = members[m];
    if (hipSetDevice(c->device) != hipSuccess) return fail(SV_E_HIP);
    if (int rc = sv_ensure_loop(c)) return fail_msg(rc, c->err.c_str());
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(SV_E_HIP);
    sv_segk &S = E->hseg[m];
    S.base = (int)nodes;
    S.npfb = c->npfb;
    S.blk0 = blocks;
    S.nblk = (c->npfb + SV_BLOCK - 1) / SV_BLOCK;
    S.volsum = c->volsum;
    S.pk = c->pk;
    sell_base[m] = sell;
    blocks += S.nblk;
    nodes += (int64_t)S.nblk * SV_BLOCK;
    sell += c->L.nnz;
    E->any_forcing |= c->pk.kol_amp != 0.0;
    E->any_mdiff |= c->pk.Mdiff != 0.0;
    E->has_mirror |= c->has_mirror;
    E->nb += c->nb;
  }
  // mirror flip codes sit above the slot in the halo keys
  if (nodes > INT32_MAX / 2)
    return fail_msg(SV_E_ARG, "the members' nodes exceed 2^30 in total");
  if (E->has_mirror && nodes > SV_FLIP_MASK)
    return fail_msg(SV_E_ARG, "members with mirror ghosts: the members' block-padded nodes "
                              "exceed the halo key's slot field");
  (void)hipSetDevice(E->device);
  E->ntot = (int)nodes;
  E->nblk = blocks;
  E->npfb = E->ntot;
  E->L.nslices = E->ntot / SV_WAVE;
  E->L.nnz = sell;
  const size_t N = (size_t)E->ntot, D = sizeof(double), I = sizeof(int);
  sv_loopl &EL = E->L;
  void **bufs[] = {(void **)&EL.off, (void **)&EL.cnt, (void **)&EL.cnt_small,
                   (void **)&EL.col, (void **)&EL.wx, (void **)&EL.wy, (void **)&EL.wl,
                   (void **)&EL.wh, (void **)&EL.gsx, (void **)&EL.gsy, (void **)&EL.lsum,
                   (void **)&EL.hsum, (void **)&EL.vol, (void **)&EL.s, (void **)&E->fx,
                   (void **)&E->U, (void **)&E->Uf, (void **)&E->reg, (void **)&E->eacc,
                   (void **)&E->GA, (void **)&E->GB, (void **)&E->partial, (void **)&E->ctrl,
                   (void **)&E->segk, (void **)&E->blk_seg};
  const size_t sizes[] = {sizeof(int64_t) * (EL.nslices + 1), I * N, I * N,
                          I * (size_t)sell, D * (size_t)sell, D * (size_t)sell,
                          D * (size_t)sell, D * (size_t)sell, D * N, D * N, D * N, D * N, D * N,
                          D * N, D * N, D * SV_US * N, D * SV_US * N, D * SV_US * N,
                          D * SV_US * N, D * E->gs * N, D * E->gs * N,
                          D * (4 * (size_t)blocks + 8), sizeof(sv_ctrl) * nmembers,
                          sizeof(sv_segk) * nmembers, I * (size_t)blocks};
  for (size_t k = 0; k < sizeof(sizes) / sizeof(sizes[0]); k++)
    if (sv_alloc(E, bufs[k], sizes[k])) return fail(SV_E_NOMEM);
  E->npartial = 4 * blocks;
  if (hipMemsetAsync(EL.off, 0, sizes[0], E->stream) != hipSuccess ||
      hipMemsetAsync(EL.cnt, 0, I * N, E->stream) != hipSuccess ||
      hipMemsetAsync(EL.cnt_small, 0, I * N, E->stream) != hipSuccess ||
      hipMemsetAsync(E->fx, 0, D * N, E->stream) != hipSuccess ||
      hipMemsetAsync(EL.s, 0, D * N, E->stream) != hipSuccess ||
      hipMemsetAsync(E->U, 0, D * SV_US * N, E->stream) != hipSuccess ||
      hipMemsetAsync(E->GA, 0, D * E->gs * N, E->stream) != hipSuccess ||
      hipMemsetAsync(E->GB, 0, D * E->gs * N, E->stream) != hipSuccess)
    return fail(SV_E_HIP);
  // per-block halos: used when every member has them (see sv_loopl)
  bool halo = true;
  int hmax = 0;
  std::vector<int64_t> hbase(nmembers + 1, 0);
  for (int m = 0; m < nmembers; m++) {
    const sv_loopl &CL = members[m]->L;
    halo = halo && CL.halo;
    if (!halo) {
      // mirror flips live in the halo keys
      if (E->has_mirror)
        return fail_msg(SV_E_ARG, "a member has mirror ghosts, so every member needs "
                                  "per-block neighbour halos (a member has none: its halo "
                                  "exceeds the LDS budget or halo_sweeps is off)");
      break;
    }
    hmax = std::max(hmax, CL.hmax);
    const int nbm = E->hseg[m].nblk;
    int64_t last = 0;
    if (hipMemcpy(&last, CL.hoff + nbm, sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess)
      return fail(SV_E_HIP);
    hbase[m + 1] = hbase[m] + last;
  }
  if (halo) {
    if (sv_alloc(E, (void **)&EL.hoff, sizeof(int64_t) * ((size_t)blocks + 1)) ||
        sv_alloc(E, (void **)&EL.hidx, I * ((size_t)hbase[nmembers] + 1)) ||
        sv_alloc(E, (void **)&EL.lcol, sizeof(uint16_t) * ((size_t)sell + 1)))
      return fail(SV_E_NOMEM);
    const int64_t tot = hbase[nmembers];
    if (hipMemcpy(EL.hoff + blocks, &tot, sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess)
      return fail(SV_E_HIP);
    EL.halo = true;
    EL.hmax = hmax;
  }
  std::vector<int> bseg(blocks);
  std::vector<int> ebidx;
  std::vector<double> ebrec;
  if (E->nb > 0) {
    ebidx.assign(N, -1);
    ebrec.reserve((size_t)SV_BREC * E->nb);
  }
  for (int m = 0; m < nmembers; m++) {
    const sv_ctx *c = members[m];
    const sv_segk &S = E->hseg[m];
    const size_t n = (size_t)S.npfb, b = (size_t)S.base;
    const int64_t sb = sell_base[m];
    for (int k = 0; k < S.nblk; k++) bseg[S.blk0 + k] = m;
    const sv_loopl &CL = c->L;
    E->member_perm.push_back(CL.hperm);
    int rc = copy_d2d(E, EL.off + b / SV_WAVE, CL.off, sizeof(int64_t) * (size_t)CL.nslices);
    if (!rc) rc = sv_dev_offset_i64(E, EL.off + b / SV_WAVE, CL.nslices, sb);
    if (!rc) rc = copy_d2d(E, EL.col + sb, CL.col, I * (size_t)CL.nnz);
    if (!rc) rc = sv_dev_offset_int(E, EL.col + sb, CL.nnz, S.base);
    const double *wsrc[4] = {CL.wx, CL.wy, CL.wl, CL.wh};
    double *wdst[4] = {EL.wx, EL.wy, EL.wl, EL.wh};
    for (int k = 0; k < 4 && !rc; k++)
      rc = copy_d2d(E, wdst[k] + sb, wsrc[k], D * (size_t)CL.nnz);
    if (!rc) rc = copy_d2d(E, EL.cnt + b, CL.cnt, I * n);
    if (!rc) rc = copy_d2d(E, EL.cnt_small + b, CL.cnt_small, I * n);
    const double *nsrc[7] = {CL.gsx, CL.gsy, CL.lsum, CL.hsum, CL.vol, CL.s, c->fx};
    double *ndst[7] = {EL.gsx, EL.gsy, EL.lsum, EL.hsum, EL.vol, EL.s, E->fx};
    for (int k = 0; k < 7 && !rc; k++) rc = copy_d2d(E, ndst[k] + b, nsrc[k], D * n);
    if (!rc) rc = copy_d2d(E, E->U + b * SV_US, c->U, D * SV_US * n);
    if (!rc) rc = copy_d2d(E, E->GA + b * E->gs, c->GA, D * E->gs * n);
    if (!rc) rc = copy_d2d(E, E->ctrl + m, c->ctrl, sizeof(sv_ctrl));
    if (!rc && c->nb > 0) {
      // boundary records: member record jb becomes jb0 + jb, FD-row slots move by base
      std::vector<int> bi(n);
      std::vector<double> br((size_t)SV_BREC * c->nb);
      if (hipMemcpy(bi.data(), CL.bidx, I * n, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(br.data(), CL.brec, D * br.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(SV_E_HIP);
      const int jb0 = (int)(ebrec.size() / SV_BREC);
      for (size_t k = 0; k < n; k++) ebidx[b + k] = bi[k] >= 0 ? jb0 + bi[k] : -1;
      for (int jb = 0; jb < c->nb; jb++)
        for (int k = 4; k < 8; k++) br[(size_t)SV_BREC * jb + k] += (double)S.base;
      ebrec.insert(ebrec.end(), br.begin(), br.end());
    }
    if (!rc && halo) {
      const int64_t hb = hbase[m], hn = hbase[m + 1] - hbase[m];
      rc = copy_d2d(E, EL.hoff + S.blk0, CL.hoff, sizeof(int64_t) * (size_t)S.nblk);
      if (!rc) rc = sv_dev_offset_i64(E, EL.hoff + S.blk0, S.nblk, hb);
      if (!rc) rc = copy_d2d(E, EL.hidx + hb, CL.hidx, I * (size_t)hn);
      if (!rc) rc = sv_dev_offset_int(E, EL.hidx + hb, hn, S.base);
      if (!rc) rc = copy_d2d(E, EL.lcol + sb, CL.lcol, sizeof(uint16_t) * (size_t)CL.nnz);
    }
    if (rc) return fail(rc);
  }
  if (E->nb > 0) {
    if (sv_alloc(E, (void **)&EL.bidx, I * N) ||
        sv_alloc(E, (void **)&EL.brec, D * ebrec.size()))
      return fail(SV_E_NOMEM);
    if (hipMemcpy(EL.bidx, ebidx.data(), I * N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(EL.brec, ebrec.data(), D * ebrec.size(), hipMemcpyHostToDevice) !=
            hipSuccess)
      return fail(SV_E_HIP);
  }
  if (hipMemcpyAsync(E->segk, E->hseg.data(), sizeof(sv_segk) * nmembers,
                     hipMemcpyHostToDevice, E->stream) != hipSuccess ||
      hipMemcpyAsync(E->blk_seg, bseg.data(), I * (size_t)blocks, hipMemcpyHostToDevice,
                     E->stream) != hipSuccess ||
      hipStreamSynchronize(E->stream) != hipSuccess)
    return fail(SV_E_HIP);
  E->sell_dirty = false;
  EL.valid = true;
  E->have_phys = true;
  E->have_state = true;
  if (int rc = sv_dev_init_partials(E)) return fail(rc);
  // launch groups: the blocks of the members of each LABFM order, with the
  // largest halo among them; the autotuner keeps one launch or one per group
  if (halo) {
    std::vector<int> orders;
    for (int m = 0; m < nmembers; m++) {
      const int o = members[m]->cfg.morder;
      if (std::find(orders.begin(), orders.end(), o) == orders.end()) orders.push_back(o);
    }
    std::sort(orders.begin(), orders.end());
    if (orders.size() >= 2) {
      for (int o : orders) {
        sv_ctx::lgroup g;
        std::vector<int> bl;
        for (int m = 0; m < nmembers; m++) {
          if (members[m]->cfg.morder != o) continue;
          const sv_segk &S = E->hseg[m];
          for (int k = 0; k < S.nblk; k++) bl.push_back(S.blk0 + k);
          g.hmax = std::max(g.hmax, members[m]->L.hmax);
        }
        g.nblk = (int)bl.size();
        if (sv_alloc(E, (void **)&g.blist, I * bl.size()) ||
            hipMemcpy(g.blist, bl.data(), I * bl.size(), hipMemcpyHostToDevice) != hipSuccess)
          return fail(SV_E_NOMEM);
        E->lgroups.push_back(g);
      }
      if (int rc = sv_dev_tune_groups(E, 5)) return fail(rc);
    }
  }
  *out = E;
  return 0;
}

int32_t sv_ensemble_size(const sv_ctx *c) { return c ? c->nseg : 0; }

}  // extern "C"

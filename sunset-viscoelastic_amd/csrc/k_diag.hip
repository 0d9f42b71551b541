// k_diag.hip -- on-device diagnostics of the current state, the reductions of
// statistics.F90:305-654:
//   SV_DIAG_MASS_ENERGY  mass_and_energy_check (:305-355)
//   SV_DIAG_CONF         conf_check (:357-442)
//   SV_DIAG_KOLMOGOROV   kolmogorov_l2norm (:552-612)
//   SV_DIAG_TG           error_TG (:614-654)
// One pass over the state records: per-block partial sums (fixed-order tree),
// then one block per segment summing its blocks in order, so results are
// bitwise reproducible and every ensemble member gets its own values.  The
// reference writes them to its statistics files; here they are returned.
#include <math.h>

#include <mutex>

#include "sv_internal.h"

namespace {

constexpr int BS = SV_BLOCK;
constexpr int DQ = 7;   // partial sums per block

struct DiagArgs {
  const int *blk_seg;
  const sv_segk *seg;
  const double *U;     // state records (ro first)
  const double *G;     // gather records (ro, u, v, cxx, cxy, cyy, ...)
  int gs;
  const double *vol;   // loop order
  const int *perm;     // loop slot -> caller's node (positions); null: identity
  const double *x, *y; // caller order
  double lambda;       // SV_DIAG_KOLMOGOROV
  double expo;         // SV_DIAG_TG: exp(-8 pi^2 t / T_char / Re)
  double *part;        // [nblk][DQ]
  // SV_DIAG_TG without a host round trip: when set, expo is computed from the
  // segment's control block (its time) and physics constants
  const sv_ctrl *tg_ctrl;
};

// exp(-8 pi^2 t / T_char / Re) of the decaying Taylor-Green vortex
// (statistics.F90:620-650), Re = rho U L / mu_total, T_char = L / U
__host__ __device__ inline double tg_decay(const sv_physk &P, double time) {
  const double Re = P.visc_total > 0 ? P.rho_char * P.U_char * P.L_char / P.visc_total : 1.0;
  const double Tc = P.L_char / P.U_char;
  return exp(-8.0 * SV_PI * SV_PI * time / Tc / Re);
}

template <int KIND>
__global__ __launch_bounds__(BS) void k_diag_part(DiagArgs a) {
  __shared__ double red[DQ][BS];
  const int bid = sv_xcd_block();
  const int sg = a.blk_seg[bid];
  const sv_segk &S = a.seg[sg];
  const int li = (bid - S.blk0) * BS + (int)threadIdx.x;
  double q[DQ] = {0, 0, 0, 0, 0, 0, 0};
  if (li < S.npfb) {
    const int i = S.base + li;
    const double dV = a.vol[i];
    const double *g = a.G + (size_t)i * a.gs;
    const double u = g[SV_G_U], v = g[SV_G_V];
    const double cxx = g[SV_G_CXX], cxy = g[SV_G_CXY], cyy = g[SV_G_CYY];
    const double czz = 1.0;   // 2D: c_zz = 1 (exp of psi_zz = 0; DI c_zz = 1)
    if (KIND == SV_DIAG_MASS_ENERGY) {
      const double ro = a.U[(size_t)i * SV_US];
      q[0] = (ro - S.pk.rho_char) * dV;
      q[1] = dV;
      q[2] = 0.5 * ro * (u * u + v * v) * dV;
    } else if (KIND == SV_DIAG_CONF) {
      q[0] = dV;
      q[1] = cxx * dV;
      q[2] = cxy * dV;
      q[3] = cyy * dV;
      q[4] = czz * dV;
      q[5] = cxy * cxy * dV;
      q[6] = (cxx + cyy + czz) * dV;
    } else if (KIND == SV_DIAG_KOLMOGOROV) {
      const double yy = a.y[a.perm ? a.perm[li] : li];
      const double s2 = sin(2.0 * yy);
      const double ue = cos(2.0 * yy);
      const double cxxe = 1.0 + 2.0 * a.lambda * a.lambda * 4.0 * s2 * s2;
      const double cxye = -a.lambda * 2.0 * s2;
      q[0] = dV * (u - ue) * (u - ue);
      q[1] = dV * v * v;
      q[2] = dV * (cxx - cxxe) * (cxx - cxxe);
      q[3] = dV * (cxy - cxye) * (cxy - cxye);
      q[4] = dV * (cyy - 1.0) * (cyy - 1.0);
      q[5] = dV;
    } else {
      const int p = a.perm ? a.perm[li] : li;
      const double xx = a.x[p], yy = a.y[p];
      const double expo = a.tg_ctrl ? tg_decay(S.pk, a.tg_ctrl[sg].time) : a.expo;
      const double ue = -expo * cos(2.0 * SV_PI * xx) * sin(2.0 * SV_PI * yy);
      const double ve = expo * sin(2.0 * SV_PI * xx) * cos(2.0 * SV_PI * yy);
      const double Uex = sqrt(ue * ue + ve * ve);
      const double Unum = sqrt(u * u + v * v);
      q[0] = (Unum - Uex) * (Unum - Uex);
      q[1] = Uex * Uex;
    }
  }
#pragma unroll
  for (int k = 0; k < DQ; k++) red[k][threadIdx.x] = q[k];
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int k = 0; k < DQ; k++) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if ((int)threadIdx.x < DQ) a.part[(size_t)bid * DQ + threadIdx.x] = red[threadIdx.x][0];
}

// one block per segment: the sums over its blocks, in block order
__global__ __launch_bounds__(256) void k_diag_sum(const sv_segk *seg, const double *part,
                                                  double *out) {
  __shared__ double red[DQ][256];
  const sv_segk &S = seg[blockIdx.x];
  double q[DQ] = {0, 0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < S.nblk; b += 256)
#pragma unroll
    for (int k = 0; k < DQ; k++) q[k] += part[(size_t)(S.blk0 + b) * DQ + k];
#pragma unroll
  for (int k = 0; k < DQ; k++) red[k][threadIdx.x] = q[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int k = 0; k < DQ; k++) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if ((int)threadIdx.x < DQ) out[(size_t)blockIdx.x * DQ + threadIdx.x] = red[threadIdx.x][0];
}

// The statistics.F90 quantities of one member from its DQ sums q (the
// reductions above); L_char^2 scales 2D volumes.
__host__ __device__ inline void diag_finish(int kind, const double *q, double L_char,
                                            double *out) {
  const double L2 = L_char * L_char;
  switch (kind) {
  case SV_DIAG_MASS_ENERGY:   // statistics.F90:309-343: mass/vol, vol, ke/vol
    out[0] = (q[0] * L2) / (q[1] * L2);
    out[1] = q[1] * L2;
    out[2] = (q[2] * L2) / (q[1] * L2);
    break;
  case SV_DIAG_CONF: {        // :365-433: means of cxx cxy cyy czz, rms cxy, mean trace
    const double V = q[0] * L2;
    for (int k = 0; k < 4; k++) out[k] = (q[1 + k] * L2) / V;
    out[4] = sqrt((q[5] * L2) / V);
    out[5] = (q[6] * L2) / V;
    break;
  }
  case SV_DIAG_KOLMOGOROV:    // :563-602: L2 of u, v, cxx, cxy, cyy
    for (int k = 0; k < 5; k++) out[k] = sqrt(q[k] / q[5]);
    break;
  default:                    // :620-650
    out[0] = sqrt(q[0] / q[1]);
    break;
  }
}

__global__ void k_diag_finish(int kind, const double *sums, int member, const sv_segk *seg,
                              double *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    diag_finish(kind, sums + (size_t)DQ * member, seg[member].pk.L_char, out);
}

// Scratch of sv_diagnostics_noalloc: the block partials ([nblk][DQ]) followed
// by the per-member sums ([nseg][DQ]), 4 MiB per device (up to ~74,000
// 256-node blocks, i.e. ~19M nodes).  The host side never passes the array's
// name to a kernel: hipGetSymbolAddress gives each device's own address once,
// kept in g_scratch below.
#define SV_DIAG_SCRATCH_DOUBLES (1 << 19)
static __device__ __align__(256) double scratch[SV_DIAG_SCRATCH_DOUBLES];

#define SV_SCRATCH_MAX_DEVICES 64
std::mutex g_scratch_mu;
double *g_scratch[SV_SCRATCH_MAX_DEVICES];
hipError_t g_scratch_err[SV_SCRATCH_MAX_DEVICES];

}  // namespace

// Look up this device's copy of `scratch` (the device must be current); once
// per device, from sv_create.  A failure is kept and reported by
// sv_diagnostics_noalloc, which is the only user.
void sv_diag_scratch_init(int device) {
  if (device < 0 || device >= SV_SCRATCH_MAX_DEVICES) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  if (e == hipSuccess) g_scratch[device] = static_cast<double *>(p);
}

static int diag_sums(sv_ctx *c, int kind, double lambda, double expo, std::vector<double> &sums) {
  double *part = nullptr, *out = nullptr;
  if (sv_alloc(c, (void **)&part, sizeof(double) * DQ * (size_t)c->nblk) ||
      sv_alloc(c, (void **)&out, sizeof(double) * DQ * (size_t)c->nseg))
    return SV_E_NOMEM;
  DiagArgs a{c->blk_seg, c->segk, c->U, c->GA, c->gs, c->L.vol,
             c->is_ensemble ? nullptr : c->L.perm, c->x, c->y, lambda, expo, part, nullptr};
  int rc = 0;
  switch (kind) {
  case SV_DIAG_MASS_ENERGY: k_diag_part<SV_DIAG_MASS_ENERGY><<<c->nblk, BS, 0, c->stream>>>(a); break;
  case SV_DIAG_CONF: k_diag_part<SV_DIAG_CONF><<<c->nblk, BS, 0, c->stream>>>(a); break;
  case SV_DIAG_KOLMOGOROV: k_diag_part<SV_DIAG_KOLMOGOROV><<<c->nblk, BS, 0, c->stream>>>(a); break;
  default: k_diag_part<SV_DIAG_TG><<<c->nblk, BS, 0, c->stream>>>(a); break;
  }
  if (hipGetLastError() != hipSuccess) rc = sv_fail(c, SV_E_HIP, "diagnostics launch failed");
  if (!rc) {
    k_diag_sum<<<c->nseg, 256, 0, c->stream>>>(c->segk, part, out);
    sums.assign((size_t)DQ * c->nseg, 0.0);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(sums.data(), out, sizeof(double) * sums.size(), hipMemcpyDeviceToHost,
                       c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      rc = sv_fail(c, SV_E_HIP, "diagnostics reduction failed");
  }
  sv_free(c, part);
  sv_free(c, out);
  return rc;
}

extern "C" {

int sv_diagnostics(sv_ctx *c, int32_t kind, int32_t member, double *out) {
  if (!c) return SV_E_ARG;
  if (!out) return sv_fail(c, SV_E_ARG, "null output");
  if (!c->have_state) return sv_fail(c, SV_E_STATE, "no state (sv_set_initial first)");
  if (kind < SV_DIAG_MASS_ENERGY || kind > SV_DIAG_TG)
    return sv_fail(c, SV_E_ARG, "unknown diagnostic %d", kind);
  if (member < 0 || member >= c->nseg)
    return sv_fail(c, SV_E_ARG, "member %d out of range [0, %d)", member, c->nseg);
  if (c->is_ensemble && (kind == SV_DIAG_KOLMOGOROV || kind == SV_DIAG_TG))
    return sv_fail(c, SV_E_ARG, "position-based diagnostics are per simulation: call them on "
                                "the member's own context");
  SV_HIP(c, hipSetDevice(c->device));
  const sv_physk &P = c->hseg[member].pk;
  double expo = 0.0;
  if (kind == SV_DIAG_TG) {
    sv_ctrl h{};
    SV_HIP(c, hipStreamSynchronize(c->stream));
    SV_HIP(c, hipMemcpy(&h, c->ctrl + member, sizeof(h), hipMemcpyDeviceToHost));
    expo = tg_decay(P, h.time);
  }
  std::vector<double> sums;
  if (int rc = diag_sums(c, kind, P.lambda, expo, sums)) return rc;
  diag_finish(kind, sums.data() + (size_t)DQ * member, P.L_char, out);
  return 0;
}

int sv_diagnostics_noalloc(sv_ctx *c, int32_t kind, int32_t member, double *out_dev,
                           void *stream) {
  if (!c) return SV_E_ARG;
  if (!out_dev) return sv_fail(c, SV_E_ARG, "null output");
  if (!c->have_state) return sv_fail(c, SV_E_STATE, "no state (sv_set_initial first)");
  if (kind < SV_DIAG_MASS_ENERGY || kind > SV_DIAG_TG)
    return sv_fail(c, SV_E_ARG, "unknown diagnostic %d", kind);
  if (member < 0 || member >= c->nseg)
    return sv_fail(c, SV_E_ARG, "member %d out of range [0, %d)", member, c->nseg);
  if (c->is_ensemble && (kind == SV_DIAG_KOLMOGOROV || kind == SV_DIAG_TG))
    return sv_fail(c, SV_E_ARG, "position-based diagnostics are per simulation: call them on "
                                "the member's own context");
  const size_t need = (size_t)DQ * ((size_t)c->nblk + (size_t)c->nseg);
  if (need > SV_DIAG_SCRATCH_DOUBLES)
    return sv_fail(c, SV_E_ARG,
                   "sv_diagnostics_noalloc: %d blocks of %d nodes and %d members need %zu "
                   "scratch doubles, above its fixed %d (use sv_diagnostics)",
                   c->nblk, SV_BLOCK, c->nseg, need, SV_DIAG_SCRATCH_DOUBLES);
  SV_HIP(c, hipSetDevice(c->device));
  double *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (c->device >= 0 && c->device < SV_SCRATCH_MAX_DEVICES) base = g_scratch[c->device];
    if (!base)
      return sv_fail(c, SV_E_HIP, "sv_diagnostics_noalloc: no scratch address for device %d "
                                  "(hipGetSymbolAddress: %s)", c->device,
                     c->device >= 0 && c->device < SV_SCRATCH_MAX_DEVICES
                         ? hipGetErrorString(g_scratch_err[c->device]) : "device out of range");
  }
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  double *part = base, *sums = base + (size_t)DQ * c->nblk;
  const sv_physk &P = c->hseg[member].pk;
  DiagArgs a{c->blk_seg, c->segk, c->U, c->GA, c->gs, c->L.vol,
             c->is_ensemble ? nullptr : c->L.perm, c->x, c->y, P.lambda, 0.0, part,
             kind == SV_DIAG_TG ? c->ctrl : nullptr};
  // every block writes its partials and every member its sums: nothing to clear
  switch (kind) {
  case SV_DIAG_MASS_ENERGY: k_diag_part<SV_DIAG_MASS_ENERGY><<<c->nblk, BS, 0, st>>>(a); break;
  case SV_DIAG_CONF: k_diag_part<SV_DIAG_CONF><<<c->nblk, BS, 0, st>>>(a); break;
  case SV_DIAG_KOLMOGOROV: k_diag_part<SV_DIAG_KOLMOGOROV><<<c->nblk, BS, 0, st>>>(a); break;
  default: k_diag_part<SV_DIAG_TG><<<c->nblk, BS, 0, st>>>(a); break;
  }
  SV_HIP(c, hipGetLastError());
  k_diag_sum<<<c->nseg, 256, 0, st>>>(c->segk, part, sums);
  SV_HIP(c, hipGetLastError());
  k_diag_finish<<<1, 64, 0, st>>>(kind, sums, member, c->segk, out_dev);
  SV_HIP(c, hipGetLastError());
  return 0;
}

}  // extern "C"

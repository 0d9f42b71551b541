"""ctypes binding of libsunset_amd.so (the C ABI in include/sunset_amd.h).

`Solver` mirrors the reference's call sequence (sunset.F90:36-106):
set_domain/set_nodes (build_domain), find_neighbours, calc_labf_weights,
calc_labf_sums, filter_coefficients, set_phys/set_initial (initial_solution),
run (the time loop).  Errors raise `SunsetError` with the library's message, as
the reference stops with a message.  There is no CPU fallback: the library must
be built (``python -c "import __graft_entry__ as g; g.build()"``).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libsunset_amd.so")

FORM_DI, FORM_LC = 0, 1
DT_PID, DT_FIXED, DT_CFL = 0, 1, 2
# "pid": set_tstep_PID (step.F90:624-682); "fixed": the given dt; "cfl": dt =
# min(dt_parabolic, dt_cfl) every step (step.F90:610-611)
DT_MODES = {"pid": DT_PID, "fixed": DT_FIXED, "cfl": DT_CFL}
OP_GRAD, OP_LAP, OP_GRAD_X, OP_GRAD_Y, OP_FILTER, OP_FILTER_TERM = range(6)

PD = C.POINTER(C.c_double)
PI32 = C.POINTER(C.c_int32)


class SunsetError(RuntimeError):
    pass


class Config(C.Structure):
    _fields_ = [("morder", C.c_int32), ("forder", C.c_int32), ("form", C.c_int32),
                ("device", C.c_int32)]


class Phys(C.Structure):
    _fields_ = [("L_char", C.c_double), ("U_char", C.c_double), ("Re", C.c_double),
                ("Wi", C.c_double), ("beta", C.c_double), ("epsPTT", C.c_double),
                ("Ma", C.c_double), ("Mdiff", C.c_double), ("grav", C.c_double * 3),
                ("driving_force", C.c_double * 3), ("kol_amp", C.c_double),
                ("kol_k", C.c_double)]


_lib = None

EXPORTS = [
    "sv_create", "sv_destroy", "sv_last_error", "sv_version", "sv_set_domain", "sv_set_nodes",
    "sv_get_ghosts", "sv_find_neighbours", "sv_calc_labf_weights", "sv_calc_labf_sums",
    "sv_filter_coefficients", "sv_get_weights", "sv_get_timing", "sv_set_option", "sv_apply",
    "sv_calc_gradient", "sv_calc_laplacian", "sv_calc_gradient_only", "sv_calc_filtered_var",
    "sv_calc_filter_term", "sv_apply_dev", "sv_set_phys", "sv_set_initial", "sv_set_control",
    "sv_get_control", "sv_calc_all_rhs", "sv_run", "sv_get_state", "sv_sync",
    "sv_run_profiled", "sv_get_stencil_stats", "sv_ensemble_create", "sv_ensemble_size",
    "sv_ensemble_get_state", "sv_ensemble_get_control", "sv_adj_setup", "sv_rhs_vjp",
    "sv_rollout_forward", "sv_rollout_vjp", "sv_adj_param_grad", "sv_adj_timing",
    "sv_diagnostics", "sv_diagnostics_noalloc", "sv_set_boundaries", "sv_set_inflow",
    "sv_calc_boundary_weights", "sv_adapt_stencils", "sv_set_state", "sv_filter_variables",
    "sv_set_state_device", "sv_get_state_device", "sv_rollout_vjp_device",
    "sv_calc_laplacian_transverse_only_on_bound",
]

KERNELS = ["dt", "stage1", "stage2", "stage3", "stage4", "filter", "reduce", "apply"]


def lib():
    """Load the library (no compute happens here, so this works without a GPU)."""
    global _lib
    if _lib is not None:
        return _lib
    p = LIB_PATH
    if not os.path.exists(p):
        raise SunsetError(f"libsunset_amd.so not built: {p} is missing "
                          "(run __graft_entry__.build())")
    L = C.CDLL(p)
    L.sv_create.argtypes = [C.POINTER(Config), C.POINTER(C.c_void_p)]
    L.sv_destroy.argtypes = [C.c_void_p]
    L.sv_destroy.restype = None
    L.sv_last_error.argtypes = [C.c_void_p]
    L.sv_last_error.restype = C.c_char_p
    L.sv_version.restype = C.c_char_p
    L.sv_set_domain.argtypes = [C.c_void_p] + [C.c_double] * 4 + [PI32, C.c_double, C.c_double]
    L.sv_set_nodes.argtypes = [C.c_void_p, C.c_int32] + [PD] * 5 + [PI32]
    L.sv_get_ghosts.argtypes = [C.c_void_p, PI32, PD, PD, PI32, PI32, C.c_int32]
    L.sv_find_neighbours.argtypes = [C.c_void_p, C.c_int32, PI32, PI32, PI32]
    L.sv_adapt_stencils.argtypes = [C.c_void_p, PD, PD, PI32]
    L.sv_calc_labf_weights.argtypes = [C.c_void_p, PI32]
    L.sv_calc_labf_sums.argtypes = [C.c_void_p]
    L.sv_filter_coefficients.argtypes = [C.c_void_p]
    L.sv_get_weights.argtypes = [C.c_void_p, C.c_int32] + [PD] * 7
    L.sv_get_timing.argtypes = [C.c_void_p, C.c_char_p, PD]
    L.sv_set_option.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
    L.sv_apply.argtypes = [C.c_void_p, C.c_int32, PD, PD]
    L.sv_calc_gradient.argtypes = [C.c_void_p, PD, PD]
    L.sv_calc_laplacian.argtypes = [C.c_void_p, PD, PD]
    L.sv_calc_laplacian_transverse_only_on_bound.argtypes = [C.c_void_p, PD, PD]
    L.sv_calc_gradient_only.argtypes = [C.c_void_p, C.c_int32, PD, PD]
    L.sv_calc_filtered_var.argtypes = [C.c_void_p, PD]
    L.sv_calc_filter_term.argtypes = [C.c_void_p, PD, PD]
    L.sv_apply_dev.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_void_p),
                               C.POINTER(C.c_void_p), C.c_void_p]
    L.sv_set_phys.argtypes = [C.c_void_p, C.POINTER(Phys)]
    L.sv_set_initial.argtypes = [C.c_void_p] + [PD] * 6
    L.sv_set_control.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int32, C.c_double,
                                 C.c_double, C.c_double, C.c_int32]
    L.sv_get_control.argtypes = [C.c_void_p, PD, PD, PI32, PD, PD, PD]
    L.sv_calc_all_rhs.argtypes = [C.c_void_p, C.c_int32] + [PD] * 6
    L.sv_set_state.argtypes = [C.c_void_p] + [PD] * 6
    L.sv_filter_variables.argtypes = [C.c_void_p] + [PD] * 6
    L.sv_run.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    L.sv_get_state.argtypes = [C.c_void_p] + [PD] * 11
    L.sv_sync.argtypes = [C.c_void_p]
    L.sv_run_profiled.argtypes = [C.c_void_p, C.c_int32, PD, PD]
    L.sv_get_stencil_stats.argtypes = [C.c_void_p, PD]
    L.sv_adj_setup.argtypes = [C.c_void_p, C.c_int32]
    L.sv_rhs_vjp.argtypes = [C.c_void_p, C.POINTER(PD), C.POINTER(PD)]
    L.sv_rollout_forward.argtypes = [C.c_void_p, C.c_int32, PD]
    L.sv_rollout_vjp.argtypes = [C.c_void_p, C.c_double, C.POINTER(PD), C.POINTER(PD)]
    L.sv_set_state_device.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.c_void_p]
    L.sv_get_state_device.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.c_void_p]
    L.sv_rollout_vjp_device.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_void_p),
                                        C.POINTER(C.c_void_p), C.c_void_p]
    L.sv_adj_param_grad.argtypes = [C.c_void_p, PD]
    L.sv_diagnostics.argtypes = [C.c_void_p, C.c_int32, C.c_int32, PD]
    L.sv_diagnostics_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                         C.c_void_p]
    L.sv_set_boundaries.argtypes = [C.c_void_p, PD, PD]
    L.sv_set_inflow.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                C.c_double]
    L.sv_calc_boundary_weights.argtypes = [C.c_void_p]
    L.sv_adj_timing.argtypes = [C.c_void_p, PD, PD]
    L.sv_ensemble_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p)]
    L.sv_ensemble_size.argtypes = [C.c_void_p]
    L.sv_ensemble_size.restype = C.c_int32
    L.sv_ensemble_get_state.argtypes = [C.c_void_p, C.c_int32] + [PD] * 11
    L.sv_ensemble_get_control.argtypes = [C.c_void_p, C.c_int32, PD, PD, C.POINTER(C.c_int32),
                                          PD, PD, PD]
    _lib = L
    return L


def _d(a):
    return None if a is None else a.ctypes.data_as(PD)


def _i(a):
    return None if a is None else a.ctypes.data_as(PI32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Solver:
    """One device context: a node set, its stencils and weights, and a state."""

    def __init__(self, morder: int = 6, forder: Optional[int] = None, form: str = "lc",
                 device: int = 0):
        L = lib()
        self.morder = morder
        self.forder = morder if forder is None else forder
        self.form = FORM_LC if form == "lc" else FORM_DI
        cfg = Config(morder, self.forder, self.form, device)
        h = C.c_void_p()
        rc = L.sv_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise SunsetError(f"sv_create failed ({rc}): {L.sv_last_error(None).decode()}")
        self.h = h
        self.npfb = 0
        self.np = 0

    def close(self):
        if getattr(self, "h", None):
            lib().sv_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        if rc != 0:
            msg = lib().sv_last_error(self.h).decode()
            raise SunsetError(f"error {rc}: {msg}")

    # ---- setup ----
    def set_nodes(self, ns):
        bc = np.array(ns.bc, dtype=np.int32)
        self._ck(lib().sv_set_domain(self.h, ns.xmin, ns.xmax, ns.ymin, ns.ymax, _i(bc),
                                     float(ns.smax), 1.0))
        self._arr = [_f64(a) for a in (ns.x, ns.y, ns.s, ns.h, ns.h_small)]
        nt = np.ascontiguousarray(ns.node_type, dtype=np.int32)
        self._ck(lib().sv_set_nodes(self.h, ns.npfb, *[_d(a) for a in self._arr], _i(nt)))
        self.npfb = ns.npfb
        self.nb = int(np.count_nonzero((nt >= 0) & (nt <= 2)))
        if getattr(ns, "rnx", None) is not None:
            self._rn = [_f64(ns.rnx), _f64(ns.rny)]
            self._ck(lib().sv_set_boundaries(self.h, _d(self._rn[0]), _d(self._rn[1])))
        npv = C.c_int32()
        self._ck(lib().sv_get_ghosts(self.h, C.byref(npv), None, None, None, None, 0))
        self.np = npv.value
        return self

    def ghosts(self):
        x = np.zeros(self.np)
        y = np.zeros(self.np)
        ng = self.np - self.npfb
        ir = np.zeros(max(ng, 1), dtype=np.int32)
        vr = np.zeros(max(ng, 1), dtype=np.int32)
        self._ck(lib().sv_get_ghosts(self.h, None, _d(x), _d(y), _i(ir), _i(vr), self.np))
        return x, y, ir[:ng] - 1, vr[:ng]

    def find_neighbours(self, nplink: int = 0, copy: bool = True):
        if not copy:
            self._ck(lib().sv_find_neighbours(self.h, nplink, None, None, None))
            return None
        cnt = np.zeros(self.npfb, dtype=np.int32)
        cs = np.zeros(self.npfb, dtype=np.int32)
        self._ck(lib().sv_find_neighbours(self.h, 0, _i(cnt), _i(cs), None))
        L = int(cnt.max()) if nplink <= 0 else nplink
        link = np.zeros((self.npfb, L), dtype=np.int32)
        self._ck(lib().sv_find_neighbours(self.h, L, None, None, _i(link)))
        return cnt, cs, link - 1

    def adapt_stencils(self):
        """adapt_stencils (labf.F90:1079-1721) after find_neighbours, as
        sunset.F90:40-52 runs it without a restart; the lists are rebuilt
        with the new h.  Returns (h, h_small, reductions [2, npfb])."""
        h = np.zeros(self.npfb)
        hs = np.zeros(self.npfb)
        red = np.zeros(2 * self.npfb, dtype=np.int32)
        self._ck(lib().sv_adapt_stencils(self.h, _d(h), _d(hs), _i(red)))
        return h, hs, red.reshape(2, -1)

    def calc_weights(self, order_map=None, persistent: Optional[bool] = None,
                     variant: Optional[int] = None, mixed: Optional[int] = None):
        """persistent=None, variant=None, mixed=None: autotuned launch
        (default); persistent True/False: fixed grid kind with 16-neighbour
        tiles; variant 0-3: fixed launch variant (bit 0 persistent, bit 1
        32-neighbour tiles); mixed 0-2: per-order launches (0) or one
        mixed-order persistent launch with 16/32-neighbour tiles."""
        if mixed is not None:
            self._ck(lib().sv_set_option(self.h, b"weights_mixed", float(mixed)))
        elif variant is not None:
            self._ck(lib().sv_set_option(self.h, b"weights_variant", float(variant)))
        elif persistent is not None:
            self._ck(lib().sv_set_option(self.h, b"weights_persistent",
                                         1.0 if persistent else 0.0))
        else:
            self._ck(lib().sv_set_option(self.h, b"weights_autotune", 1.0))
        om = None if order_map is None else np.ascontiguousarray(order_map, dtype=np.int32)
        self._ck(lib().sv_calc_labf_weights(self.h, _i(om)))

    def calc_boundary_weights(self):
        """calc_boundary_weights (labf.F90:505-1039); a no-op without boundary nodes."""
        self._ck(lib().sv_calc_boundary_weights(self.h))

    def calc_sums(self):
        self._ck(lib().sv_calc_labf_sums(self.h))

    def filter_coefficients(self):
        self._ck(lib().sv_filter_coefficients(self.h))

    def setup(self, ns, order_map=None, adapt: bool = False):
        self.set_nodes(ns)
        self.find_neighbours(copy=False)
        if adapt:
            self.adapt_stencils()
        self.calc_weights(order_map)
        self.calc_boundary_weights()
        self.calc_sums()
        self.filter_coefficients()
        return self

    def set_inflow(self, inflow_type=0, u_control=0, u_start=0.0, u_end=1.0, ramp_time=1.0):
        """The inflow line of control.in (load_data.F90:152): type 0 soft,
        1 hard, 2 inflow-outflow; u_control 1 ramps the inflow velocity."""
        self._ck(lib().sv_set_inflow(self.h, inflow_type, u_control, u_start, u_end, ramp_time))

    def set_option(self, key: str, value: float):
        """sv_set_option (e.g. "halo_sweeps" 0/1 before set_initial)."""
        self._ck(lib().sv_set_option(self.h, key.encode(), float(value)))

    def timing(self, what: str) -> float:
        v = C.c_double()
        self._ck(lib().sv_get_timing(self.h, what.encode(), C.byref(v)))
        return v.value

    def weights(self, nplink: int):
        n = self.npfb
        wg = np.zeros((n, nplink, 2))
        wl = np.zeros((n, nplink))
        wh = np.zeros((n, nplink))
        gs = np.zeros((n, 2))
        ls, hs, vol = np.zeros(n), np.zeros(n), np.zeros(n)
        self._ck(lib().sv_get_weights(self.h, nplink, _d(wg), _d(wl), _d(wh), _d(gs), _d(ls),
                                      _d(hs), _d(vol)))
        return dict(w_grad=wg, w_lap=wl, w_hyp=wh, gsum=gs, lsum=ls, hsum=hs, vol=vol)

    # ---- drop-in operators (phi sized np) ----
    def gradient(self, phi):
        phi = _f64(phi)
        out = np.zeros(2 * self.npfb)
        self._ck(lib().sv_calc_gradient(self.h, _d(phi), _d(out)))
        return out[: self.npfb], out[self.npfb:]

    def laplacian(self, phi):
        phi = _f64(phi)
        out = np.zeros(self.npfb)
        self._ck(lib().sv_calc_laplacian(self.h, _d(phi), _d(out)))
        return out

    def laplacian_transverse(self, phi):
        """calc_laplacian_transverse_only_on_bound (derivatives.F90:306-358):
        the Laplacian, tangential only on boundary nodes."""
        phi = _f64(phi)
        out = np.zeros(self.npfb)
        self._ck(lib().sv_calc_laplacian_transverse_only_on_bound(self.h, _d(phi), _d(out)))
        return out

    def gradient_only(self, dim, phi):
        phi = _f64(phi)
        out = np.zeros(self.npfb)
        self._ck(lib().sv_calc_gradient_only(self.h, dim, _d(phi), _d(out)))
        return out

    def filter_term(self, phi):
        phi = _f64(phi)
        out = np.zeros(self.npfb)
        self._ck(lib().sv_calc_filter_term(self.h, _d(phi), _d(out)))
        return out

    def filtered_var(self, phi):
        phi = np.array(phi, dtype=np.float64)
        self._ck(lib().sv_calc_filtered_var(self.h, _d(phi)))
        return phi[: self.npfb]

    # ---- physics ----
    def set_phys(self, Re=1.0, Wi=1.0, beta=0.5, eps=0.0, Ma=0.1, Mdiff=0.0, kol_amp=0.0,
                 kol_k=0.0, U_char=1.0, L_char=1.0, grav=(0, 0, 0), drive=(0, 0, 0)):
        p = Phys()
        p.L_char, p.U_char, p.Re, p.Wi, p.beta, p.epsPTT, p.Ma, p.Mdiff = (
            L_char, U_char, Re, Wi, beta, eps, Ma, Mdiff)
        for k in range(3):
            p.grav[k] = grav[k]
            p.driving_force[k] = drive[k]
        p.kol_amp, p.kol_k = kol_amp, kol_k
        self._ck(lib().sv_set_phys(self.h, C.byref(p)))
        self.phys = dict(Re=Re, Wi=Wi, beta=beta, eps=eps, Ma=Ma, Mdiff=Mdiff, kol_amp=kol_amp,
                         kol_k=kol_k, U_char=U_char, L_char=L_char, grav=tuple(grav),
                         drive=tuple(drive))

    def set_initial(self, ic, dt=None, itime=0, dt_mode="pid", time=0.0,
                    emax=(1e-5, 1e-5, 1e-5)):
        arrs = [_f64(ic[k]) for k in ("ro", "u", "v", "cxx", "cxy", "cyy")]
        self._ck(lib().sv_set_initial(self.h, *[_d(a) for a in arrs]))
        self._ck(lib().sv_set_control(self.h, 1.0e-6 if dt is None else dt, time, itime,
                                      emax[0], emax[1], emax[2],
                                      DT_MODES[dt_mode]))

    def run(self, nsteps: int, graph: bool = False):
        self._ck(lib().sv_run(self.h, nsteps, 1 if graph else 0))

    def run_profiled(self, nsteps: int):
        """Eager steps with HIP events between launches: (ms per kernel, total ms)."""
        k = np.zeros(len(KERNELS))
        tot = C.c_double()
        self._ck(lib().sv_run_profiled(self.h, nsteps, _d(k), C.byref(tot)))
        return dict(zip(KERNELS, k)), tot.value

    def stencil_stats(self):
        st = np.zeros(6)
        self._ck(lib().sv_get_stencil_stats(self.h, _d(st)))
        return dict(K_mean=st[0], Ks_mean=st[1], K_max=int(st[2]), sell_entries=int(st[3]),
                    ghosts=int(st[4]), halo_per_node=st[5])

    def rhs(self, iRKstep=1):
        outs = [np.zeros(self.npfb) for _ in range(6)]
        self._ck(lib().sv_calc_all_rhs(self.h, iRKstep, *[_d(a) for a in outs]))
        return dict(zip(("ro", "rou", "rov", "xx", "xy", "yy"), outs))

    # ---- reverse mode (DI, fixed dt) ----
    STATE_KEYS = ("ro", "rou", "rov", "cxx", "cxy", "cyy")

    def _ptrs(self, d, alloc=False):
        if d is None:
            return None, []
        arrs = [np.zeros(self.npfb) if alloc else _f64(d[k]) for k in self.STATE_KEYS]
        return (PD * 6)(*[_d(a) for a in arrs]), arrs

    def adj_setup(self, checkpoint_every: int = 20):
        self._ck(lib().sv_adj_setup(self.h, checkpoint_every))

    def rhs_vjp(self, rbar):
        """(d calc_all_rhs / d state)^T rbar; rbar keyed ro, rou, rov, cxx, cxy, cyy."""
        pin, keep = self._ptrs(rbar)
        pout, out = self._ptrs(None, alloc=False)[0], [np.zeros(self.npfb) for _ in range(6)]
        pout = (PD * 6)(*[_d(a) for a in out])
        self._ck(lib().sv_rhs_vjp(self.h, pin, pout))
        return dict(zip(self.STATE_KEYS, out))

    def rollout_forward(self, nsteps: int) -> float:
        L = C.c_double()
        self._ck(lib().sv_rollout_forward(self.h, nsteps, C.byref(L)))
        return L.value

    def rollout_vjp(self, loss_bar: float = 1.0, ubar_final=None):
        pin, keep = self._ptrs(ubar_final)
        out = [np.zeros(self.npfb) for _ in range(6)]
        pout = (PD * 6)(*[_d(a) for a in out])
        self._ck(lib().sv_rollout_vjp(self.h, loss_bar, pin, pout))
        return dict(zip(self.STATE_KEYS, out))

    # ---- device-pointer forms (integers: device addresses, e.g. a torch
    # tensor's data_ptr(); the library must share the caller's HIP runtime) ----
    def set_state_device(self, ptrs, stream: int = 0):
        """The registers (ro, rou, rov, xx, xy, yy), six device arrays of npfb
        float64 in the caller's node order; after set_initial."""
        arr = (C.c_void_p * 6)(*[C.c_void_p(int(p)) for p in ptrs])
        self._ck(lib().sv_set_state_device(self.h, arr, C.c_void_p(stream or None)))

    def get_state_device(self, ptrs, stream: int = 0):
        arr = (C.c_void_p * 6)(*[C.c_void_p(int(p)) if p else None for p in ptrs])
        self._ck(lib().sv_get_state_device(self.h, arr, C.c_void_p(stream or None)))

    def rollout_vjp_device(self, grad_ptrs, loss_bar: float = 1.0, ubar_ptrs=None,
                           stream: int = 0):
        out = (C.c_void_p * 6)(*[C.c_void_p(int(p)) for p in grad_ptrs])
        inp = None
        if ubar_ptrs is not None:
            inp = (C.c_void_p * 6)(*[C.c_void_p(int(p)) if p else None for p in ubar_ptrs])
        self._ck(lib().sv_rollout_vjp_device(self.h, loss_bar, inp, out,
                                              C.c_void_p(stream or None)))

    def diagnostics(self, kind: str):
        """On-device statistics.F90 reductions of the current state:
        "mass_energy" (mass/vol, vol, ke/vol), "conf" (means of cxx, cxy,
        cyy, czz, rms cxy, mean trace), "kolmogorov" (L2 errors of u, v, cxx,
        cxy, cyy), "tg" (relative L2 error of |u|)."""
        return _diagnostics(self.h, self._ck, kind, 0)

    def adj_param_grad(self):
        """d/d(Wi, beta, eps) of the quantity of the last rollout_vjp (or of
        <rbar, rhs> for the last rhs_vjp)."""
        g = np.zeros(3)
        self._ck(lib().sv_adj_param_grad(self.h, _d(g)))
        return dict(Wi=g[0], beta=g[1], eps=g[2])

    def adj_timing(self):
        f, b = C.c_double(), C.c_double()
        self._ck(lib().sv_adj_timing(self.h, C.byref(f), C.byref(b)))
        return f.value, b.value

    def set_state(self, U, dt=None, itime=0, time=0.0):
        """Set (ro, rou, rov, cxx, cxy, cyy) directly (DI) with a fixed dt: the
        records of set_initial, then the registers exactly as given."""
        ro = _f64(U["ro"])
        ic = dict(ro=ro, u=_f64(U["rou"]) / ro, v=_f64(U["rov"]) / ro, cxx=U["cxx"],
                  cxy=U["cxy"], cyy=U["cyy"])
        self.set_initial(ic, dt=dt, itime=itime, dt_mode="fixed", time=time)
        arrs = [_f64(U[k]) for k in ("ro", "rou", "rov", "cxx", "cxy", "cyy")]
        self._ck(lib().sv_set_state(self.h, *[_d(a) for a in arrs]))

    def set_state_control(self, dt, itime=0, time=0.0, emax=(1e-5, 1e-5, 1e-5)):
        """A fixed dt, time and itime for the current state (restart-style)."""
        self._ck(lib().sv_set_control(self.h, dt, time, itime, emax[0], emax[1], emax[2],
                                      DT_FIXED))

    def fields(self):
        n = self.npfb
        keys = ("ro", "rou", "rov", "u", "v", "pxx", "pxy", "pyy", "cxx", "cxy", "cyy")
        out = {k: np.zeros(n) for k in keys}
        self._ck(lib().sv_get_state(self.h, *[_d(out[k]) for k in keys]))
        dt, tm, dc, dp = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        it = C.c_int32()
        em = np.zeros(3)
        self._ck(lib().sv_get_control(self.h, C.byref(dt), C.byref(tm), C.byref(it), _d(em),
                                      C.byref(dc), C.byref(dp)))
        out.update(dt=dt.value, time=tm.value, itime=it.value, emax_np1=em[0], emax_n=em[1],
                   emax_nm1=em[2], dt_cfl=dc.value, dt_par=dp.value)
        return out


_FIELD_KEYS = ("ro", "rou", "rov", "u", "v", "pxx", "pxy", "pyy", "cxx", "cxy", "cyy")


DIAG = {"mass_energy": (0, 3), "conf": (1, 6), "kolmogorov": (2, 5), "tg": (3, 1)}


def _diagnostics(h, ck, kind, member):
    k, n = DIAG[kind]
    out = np.zeros(n)
    ck(lib().sv_diagnostics(h, k, member, _d(out)))
    return out


class Ensemble:
    """Several Solver contexts (each fully set up, with an initial state)
    advanced together: one launch per kernel and one hipGraph per step for all
    members.  The members' own contexts are left unchanged."""

    def __init__(self, members):
        L = lib()
        members = list(members)
        if not members:
            raise SunsetError("an ensemble needs at least one member")
        arr = (C.c_void_p * len(members))(*[m.h.value for m in members])
        h = C.c_void_p()
        rc = L.sv_ensemble_create(arr, len(members), C.byref(h))
        if rc != 0:
            raise SunsetError(f"sv_ensemble_create failed ({rc}): "
                              f"{L.sv_last_error(None).decode()}")
        self.h = h
        self.sizes = [m.npfb for m in members]

    def __len__(self):
        """Members, as the library counts them (sv_ensemble_size)."""
        return int(lib().sv_ensemble_size(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().sv_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        if rc != 0:
            raise SunsetError(f"error {rc}: {lib().sv_last_error(self.h).decode()}")

    def diagnostics(self, member: int, kind: str):
        """Member's "mass_energy" or "conf" reductions (see Solver.diagnostics)."""
        return _diagnostics(self.h, self._ck, kind, member)

    def run(self, nsteps: int, graph: bool = False):
        self._ck(lib().sv_run(self.h, nsteps, 1 if graph else 0))

    def run_profiled(self, nsteps: int):
        k = np.zeros(len(KERNELS))
        tot = C.c_double()
        self._ck(lib().sv_run_profiled(self.h, nsteps, _d(k), C.byref(tot)))
        return dict(zip(KERNELS, k)), tot.value

    def step_ms(self) -> float:
        return self.timing("step")

    def timing(self, what: str) -> float:
        """sv_get_timing: "step", "ensemble_groups" (the launch mode in use: 0
        one launch over all members, 1 one launch per LABFM-order group),
        "ensemble_groups_tune_0|1" (the autotuner's sample times, ms)."""
        v = C.c_double()
        self._ck(lib().sv_get_timing(self.h, what.encode(), C.byref(v)))
        return v.value

    def set_option(self, key: str, value: float):
        """sv_set_option; "ensemble_groups": -1 retune, 0 / 1 fix the launch mode."""
        self._ck(lib().sv_set_option(self.h, key.encode(), float(value)))

    def fields(self, member: int):
        n = self.sizes[member]
        out = {k: np.zeros(n) for k in _FIELD_KEYS}
        self._ck(lib().sv_ensemble_get_state(self.h, member, *[_d(out[k]) for k in _FIELD_KEYS]))
        dt, tm, dc, dp = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        it = C.c_int32()
        em = np.zeros(3)
        self._ck(lib().sv_ensemble_get_control(self.h, member, C.byref(dt), C.byref(tm),
                                               C.byref(it), _d(em), C.byref(dc), C.byref(dp)))
        out.update(dt=dt.value, time=tm.value, itime=it.value, emax_np1=em[0], emax_n=em[1],
                   emax_nm1=em[2], dt_cfl=dc.value, dt_par=dp.value)
        return out

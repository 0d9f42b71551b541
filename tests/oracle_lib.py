"""ctypes view of the CPU oracle (oracle/_build/libsv_oracle.so).

Test infrastructure: the oracle is the checker for the HIP path.  Only tests,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg use it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "libsv_oracle.so")

PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)


class Domain(C.Structure):
    _fields_ = [("npfb", C.c_int), ("np", C.c_int), ("nplink", C.c_int), ("cap", C.c_int),
                ("x", PD), ("y", PD), ("h", PD), ("hs", PD), ("s", PD),
                ("node_type", PI), ("irel", PI), ("vrel", PI),
                ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double),
                ("ymax", C.c_double), ("smax", C.c_double), ("h0", C.c_double),
                ("h2", C.c_double), ("bc", C.c_int * 4), ("morder", C.c_int),
                ("forder", C.c_int), ("ij_count", PI), ("ij_count_small", PI),
                ("ij_link", PI), ("w_grad", PD), ("w_lap", PD), ("w_hyp", PD),
                ("gsum", PD), ("lsum", PD), ("hsum", PD), ("vol", PD),
                ("L_char", C.c_double), ("nb", C.c_int), ("blist", PI), ("fdp", PI),
                ("rnx", PD), ("rny", PD), ("wbt", PD), ("wbtsum", PD)]


class Phys(C.Structure):
    _fields_ = [("L_char", C.c_double), ("U_char", C.c_double), ("Re", C.c_double),
                ("Wi", C.c_double), ("beta", C.c_double), ("epsPTT", C.c_double),
                ("Ma", C.c_double), ("Mdiff", C.c_double), ("grav", C.c_double * 3),
                ("drive", C.c_double * 3), ("kol_amp", C.c_double), ("kol_k", C.c_double),
                ("form", C.c_int), ("dt_mode", C.c_int), ("rho_char", C.c_double),
                ("visc_total", C.c_double), ("visc_solvent", C.c_double),
                ("visc_polymeric", C.c_double), ("lam", C.c_double), ("csq", C.c_double),
                ("inflow_type", C.c_int), ("u_control", C.c_int), ("u_start", C.c_double),
                ("u_end", C.c_double), ("ramp_time", C.c_double)]


class State(C.Structure):
    _fields_ = [("np", C.c_int), ("ro", PD), ("rou", PD), ("rov", PD), ("u", PD), ("v", PD),
                ("pxx", PD), ("pxy", PD), ("pyy", PD), ("cxx", PD), ("cxy", PD), ("cyy", PD),
                ("dt", C.c_double), ("time", C.c_double), ("dt_cfl", C.c_double),
                ("dt_par", C.c_double), ("emax_np1", C.c_double), ("emax_n", C.c_double),
                ("emax_nm1", C.c_double), ("itime", C.c_int), ("u_in", PD)]


class Tangent(C.Structure):
    _fields_ = [("np", C.c_int), ("ro", PD), ("rou", PD), ("rov", PD), ("u", PD), ("v", PD),
                ("cxx", PD), ("cxy", PD), ("cyy", PD), ("dWi", C.c_double),
                ("dbeta", C.c_double), ("deps", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.or_domain_init.argtypes = [C.POINTER(Domain), C.c_int, PD, PD, PD, PD, PD, PI,
                                     C.c_double, C.c_double, C.c_double, C.c_double, PI,
                                     C.c_double, C.c_int, C.c_int, C.c_double]
        L.or_wab.restype = C.c_double
        L.or_wab.argtypes = [C.c_double, C.c_double]
        L.or_monomials.argtypes = [C.c_int, C.c_double, C.c_double, PD]
        L.or_abfs.argtypes = [C.c_int, C.c_double, C.c_double, PD]
        L.or_svd_solve.argtypes = [C.c_int, PD, PD]
        L.or_eigens2.argtypes = [C.c_double, C.c_double, C.c_double, PD, PD]
        L.or_c_from_psi.argtypes = [C.c_double] * 3 + [PD] * 3
        L.or_psi_from_c.argtypes = [C.c_double] * 3 + [PD] * 3
        L.or_calc_gradient.argtypes = [C.POINTER(Domain), PD, PD, PD]
        L.or_calc_laplacian.argtypes = [C.POINTER(Domain), PD, PD]
        L.or_calc_laplacian_transverse.argtypes = [C.POINTER(Domain), PD, PD]
        L.or_calc_gradient_only.argtypes = [C.POINTER(Domain), C.c_int, PD, PD]
        L.or_calc_filtered_var.argtypes = [C.POINTER(Domain), PD]
        L.or_calc_filter_term.argtypes = [C.POINTER(Domain), PD, PD]
        L.or_calc_all_rhs.argtypes = [C.POINTER(Domain), C.POINTER(Phys), C.POINTER(State),
                                      C.c_int] + [PD] * 6
        L.or_set_initial.argtypes = [C.POINTER(Domain), C.POINTER(Phys), C.POINTER(State)] + [PD] * 6
        L.or_run.argtypes = [C.POINTER(Domain), C.POINTER(Phys), C.POINTER(State), C.c_int]
        L.or_calc_labf_weights_ordermap.argtypes = [C.POINTER(Domain), PI, C.c_int]
        L.or_tangent_alloc.argtypes = [C.POINTER(Tangent), C.c_int]
        L.or_tangent_free.argtypes = [C.POINTER(Tangent)]
        L.or_tangent_set.argtypes = [C.POINTER(Domain), C.POINTER(State),
                                     C.POINTER(Tangent)] + [PD] * 6
        L.or_calc_all_rhs_di_tlm.argtypes = [C.POINTER(Domain), C.POINTER(Phys),
                                             C.POINTER(State), C.POINTER(Tangent), C.POINTER(PD)]
        L.or_rollout_tlm.argtypes = [C.POINTER(Domain), C.POINTER(Phys), C.POINTER(State),
                                     C.POINTER(Tangent), C.c_int, PD, PD]
        L.or_loss_trace.argtypes = [C.POINTER(Domain), C.POINTER(State)]
        L.or_loss_trace.restype = C.c_double
        L.or_diag_mass_energy.argtypes = [C.POINTER(Domain), C.POINTER(Phys), C.POINTER(State),
                                          PD]
        L.or_diag_conf.argtypes = [C.POINTER(Domain), C.POINTER(State), PD]
        L.or_diag_kolmogorov.argtypes = [C.POINTER(Domain), C.POINTER(Phys), C.POINTER(State),
                                         PD]
        L.or_diag_tg.argtypes = [C.POINTER(Domain), C.POINTER(Phys), C.POINTER(State)]
        L.or_diag_tg.restype = C.c_double
        L.or_set_boundaries.argtypes = [C.POINTER(Domain), PD, PD]
        L.or_boundary_weights.argtypes = [C.POINTER(Domain)]
        L.or_adapt_stencils.argtypes = [C.POINTER(Domain), PI]
        L.or_svd_solve_gr.argtypes = [C.c_int, PD, PD]
        L.or_svd_gr.argtypes = [C.c_int, PD, PD, PD, PD]
        L.or_adapt_check.argtypes = [C.POINTER(Domain), PD, C.c_int, PD, PI]
        L.or_set_wab_pow4.argtypes = [C.c_int]
        L.or_set_stencil.argtypes = [C.POINTER(Domain), C.c_int, PI, PI, PI]
        L.or_set_weights.argtypes = [C.POINTER(Domain)] + [PD] * 7
        L.or_weights_subset.argtypes = [C.POINTER(Domain), PI, C.c_int, PI]
        L.or_set_num_threads.argtypes = [C.c_int]
        L.or_set_plant.argtypes = [C.c_int]
        L.or_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _pd(a):
    return a.ctypes.data_as(PD)


def _pi(a):
    return a.ctypes.data_as(PI)


def _view(ptr, n, dtype=np.float64):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).copy()


class wab_pow4:
    """Context: the oracle's Wendland (1 - q/2)^4 as libm's pow(t, 4.0)
    (mode 0, the reference's) or correctly rounded (mode 1, the device's)."""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        lib().or_set_wab_pow4(self.mode)
        return self

    def __exit__(self, *exc):
        lib().or_set_wab_pow4(0)


class OracleSim:
    """The oracle's domain + state for one node set (0-based indices)."""

    def __init__(self, ns, morder, forder=None, L_char=1.0, nplink=None):
        L = lib()
        self.ns = ns
        self.morder = morder
        self.forder = forder if forder is not None else morder
        self.d = Domain()
        self._keep = [np.ascontiguousarray(a, dtype=np.float64)
                      for a in (ns.x, ns.y, ns.s, ns.h, ns.h_small)]
        nt = np.ascontiguousarray(ns.node_type, dtype=np.int32)
        bc = np.array(ns.bc, dtype=np.int32)
        rc = L.or_domain_init(C.byref(self.d), ns.npfb, *[_pd(a) for a in self._keep], _pi(nt),
                              ns.xmin, ns.xmax, ns.ymin, ns.ymax, _pi(bc), ns.smax, morder,
                              self.forder, L_char)
        assert rc == 0
        from sunset_amd.nodes import nplink_default
        self.nplink = nplink if nplink is not None else nplink_default(self.forder)
        self.st = None

    def __del__(self):
        try:
            lib().or_domain_free(C.byref(self.d))
            if self.st is not None:
                lib().or_state_free(C.byref(self.st))
        except Exception:
            pass

    # ---- setup ----
    def ghosts(self):
        rc = lib().or_create_ghosts(C.byref(self.d))
        assert rc == 0, rc
        if getattr(self.ns, "rnx", None) is not None:
            self._rn = [np.ascontiguousarray(self.ns.rnx, dtype=np.float64),
                        np.ascontiguousarray(self.ns.rny, dtype=np.float64)]
            rc = lib().or_set_boundaries(C.byref(self.d), _pd(self._rn[0]), _pd(self._rn[1]))
            assert rc == 0, "malformed boundary layout"

    @property
    def nb(self):
        return self.d.nb

    def neighbours(self):
        rc = lib().or_find_neighbours(C.byref(self.d), self.nplink)
        assert rc == 0, f"neighbour list overflow (nplink={self.nplink})"

    def weights(self):
        assert lib().or_calc_labf_weights(C.byref(self.d)) == 0
        if self.d.nb:
            assert lib().or_boundary_weights(C.byref(self.d)) == 0
        lib().or_calc_labf_sums(C.byref(self.d))
        lib().or_filter_coefficients(C.byref(self.d))

    def weights_ordermap(self, order_map, nthreads=0):
        """Weights with a per-node LABFM order (derivative and filter system)."""
        om = np.ascontiguousarray(order_map, dtype=np.int32)
        assert om.shape == (self.d.npfb,)
        self._om = om
        assert lib().or_calc_labf_weights_ordermap(C.byref(self.d), _pi(om), nthreads) == 0
        lib().or_calc_labf_sums(C.byref(self.d))
        lib().or_filter_coefficients(C.byref(self.d))

    def setup(self):
        self.ghosts()
        self.neighbours()
        self.weights()
        return self

    # ---- checker set-up for full-size node sets ----
    def load_lists(self, cnt, cnt_small, link):
        """Neighbour lists computed elsewhere (0-based link [npfb, nplink])."""
        link = np.ascontiguousarray(link, dtype=np.int32)
        self._lists = [np.ascontiguousarray(cnt, dtype=np.int32),
                       np.ascontiguousarray(cnt_small, dtype=np.int32), link]
        self.nplink = link.shape[1]
        rc = lib().or_set_stencil(C.byref(self.d), self.nplink, *[_pi(a) for a in self._lists])
        assert rc == 0, rc

    def load_weights(self, W):
        """Finished weights (after sums and filter coefficients) in the
        reference layout, as Solver.weights(nplink) returns them."""
        arrs = [np.ascontiguousarray(W[k], dtype=np.float64)
                for k in ("w_grad", "w_lap", "w_hyp", "gsum", "lsum", "hsum", "vol")]
        assert arrs[1].shape == (self.d.npfb, self.d.nplink)
        assert lib().or_set_weights(C.byref(self.d), *[_pd(a) for a in arrs]) == 0

    def adapt_stencils(self):
        """adapt_stencils + adapt_stencils_main (labf.F90:1079-1721) on the
        current lists: h, h_small and the lists change; returns the accepted
        2% reductions per node of the two passes ([2, npfb])."""
        red = np.zeros(2 * self.d.npfb, dtype=np.int32)
        assert lib().or_adapt_stencils(C.byref(self.d), _pi(red)) == 0, "SVD not converging"
        return red.reshape(2, -1)

    def adapt_check(self, h_sel, pass_):
        """The reference's acceptance checks of stencils h_sel (pass 1: h,
        pass 2: h_small) on the current lists: (residual measure / tolerance,
        first failing amplification wavenumber or 0) per node."""
        h_sel = np.ascontiguousarray(h_sel, dtype=np.float64)
        ratio = np.zeros(self.d.npfb)
        amp = np.zeros(self.d.npfb, dtype=np.int32)
        assert lib().or_adapt_check(C.byref(self.d), _pd(h_sel), pass_, _pd(ratio),
                                    _pi(amp)) == 0
        return ratio, amp

    def h_hs(self):
        n = self.d.npfb
        return _view(self.d.h, n).copy(), _view(self.d.hs, n).copy()

    def weights_subset(self, nodes, order_map=None):
        """The weight solve (derivative and filter systems, volumes) of the
        listed nodes only; the other nodes' weights stay zero."""
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        om = None if order_map is None else np.ascontiguousarray(order_map, dtype=np.int32)
        self._om_subset = om
        rc = lib().or_weights_subset(C.byref(self.d), _pi(nodes), nodes.size,
                                     None if om is None else _pi(om))
        assert rc == 0, rc

    # ---- array views ----
    @property
    def np_(self):
        return self.d.np

    def xy(self):
        return _view(self.d.x, self.d.np), _view(self.d.y, self.d.np)

    def irel(self):
        return _view(self.d.irel, self.d.np, np.int32)

    def lists(self):
        n, L = self.d.npfb, self.d.nplink
        cnt = _view(self.d.ij_count, n, np.int32)
        cs = _view(self.d.ij_count_small, n, np.int32)
        link = _view(self.d.ij_link, n * L, np.int32).reshape(n, L)
        return cnt, cs, link

    def weight_arrays(self):
        n, L = self.d.npfb, self.d.nplink
        wg = _view(self.d.w_grad, n * L * 2).reshape(n, L, 2)
        wl = _view(self.d.w_lap, n * L).reshape(n, L)
        wh = _view(self.d.w_hyp, n * L).reshape(n, L)
        gs = _view(self.d.gsum, 2 * n).reshape(n, 2)
        return dict(w_grad=wg, w_lap=wl, w_hyp=wh, gsum=gs, lsum=_view(self.d.lsum, n),
                    hsum=_view(self.d.hsum, n), vol=_view(self.d.vol, n))

    # ---- operators on a field of length np ----
    def gradient(self, phi):
        n = self.d.npfb
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        gx, gy = np.zeros(n), np.zeros(n)
        lib().or_calc_gradient(C.byref(self.d), _pd(phi), _pd(gx), _pd(gy))
        return gx, gy

    def laplacian(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        out = np.zeros(self.d.npfb)
        lib().or_calc_laplacian(C.byref(self.d), _pd(phi), _pd(out))
        return out

    def laplacian_transverse(self, phi):
        """calc_laplacian_transverse_only_on_bound (derivatives.F90:306-358)."""
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        out = np.zeros(self.d.npfb)
        lib().or_calc_laplacian_transverse(C.byref(self.d), _pd(phi), _pd(out))
        return out

    def gradient_only(self, dim, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        out = np.zeros(self.d.npfb)
        lib().or_calc_gradient_only(C.byref(self.d), dim, _pd(phi), _pd(out))
        return out

    def filter_term(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        out = np.zeros(self.d.npfb)
        lib().or_calc_filter_term(C.byref(self.d), _pd(phi), _pd(out))
        return out

    def filtered_var(self, phi):
        phi = np.array(phi, dtype=np.float64)
        lib().or_calc_filtered_var(C.byref(self.d), _pd(phi))
        return phi[: self.d.npfb]

    def extend(self, f):
        """Field on real nodes -> field on real+ghost nodes (periodic copy)."""
        irel = self.irel()
        out = np.empty(self.d.np)
        out[: self.d.npfb] = f
        out[self.d.npfb:] = f[irel[self.d.npfb:]]
        return out

    # ---- physics ----
    def make_phys(self, form="lc", dt_mode="pid", Re=1.0, Wi=1.0, beta=0.5, eps=0.0, Ma=0.1,
                  Mdiff=0.0, kol_amp=0.0, kol_k=0.0, U_char=1.0, L_char=1.0, drive=(0.0, 0.0),
                  inflow_type=0, u_control=0, u_start=0.0, u_end=1.0, ramp_time=1.0):
        p = Phys()
        p.drive[0], p.drive[1] = drive
        p.inflow_type, p.u_control = inflow_type, u_control
        p.u_start, p.u_end, p.ramp_time = u_start, u_end, ramp_time
        p.L_char, p.U_char, p.Re, p.Wi, p.beta, p.epsPTT, p.Ma, p.Mdiff = (
            L_char, U_char, Re, Wi, beta, eps, Ma, Mdiff)
        p.kol_amp, p.kol_k = kol_amp, kol_k
        p.form = 1 if form == "lc" else 0
        p.dt_mode = {"pid": 0, "fixed": 1, "cfl": 2}[dt_mode]
        lib().or_phys_derive(C.byref(p))
        self.p = p
        return p

    def set_initial(self, ic, dt=None, itime=0):
        if self.st is None:
            self.st = State()
            assert lib().or_state_alloc(C.byref(self.st), self.d.np) == 0
        arrs = [np.ascontiguousarray(ic[k], dtype=np.float64)
                for k in ("ro", "u", "v", "cxx", "cxy", "cyy")]
        lib().or_set_initial(C.byref(self.d), C.byref(self.p), C.byref(self.st),
                             *[_pd(a) for a in arrs])
        if dt is not None:
            self.st.dt = dt
        self.st.itime = itime

    def run(self, nsteps):
        lib().or_run(C.byref(self.d), C.byref(self.p), C.byref(self.st), nsteps)

    def rhs(self, iRKstep=1):
        n = self.d.npfb
        outs = [np.zeros(n) for _ in range(6)]
        lib().or_calc_all_rhs(C.byref(self.d), C.byref(self.p), C.byref(self.st), iRKstep,
                              *[_pd(a) for a in outs])
        return dict(zip(("ro", "rou", "rov", "xx", "xy", "yy"), outs))

    def fields(self):
        n = self.d.npfb
        st = self.st
        out = {k: _view(getattr(st, k), n) for k in
               ("ro", "rou", "rov", "u", "v", "pxx", "pxy", "pyy", "cxx", "cxy", "cyy")}
        out.update(dt=st.dt, time=st.time, emax_np1=st.emax_np1, emax_n=st.emax_n,
                   emax_nm1=st.emax_nm1, dt_cfl=st.dt_cfl, dt_par=st.dt_par, itime=st.itime)
        return out

    # ---- tangent-linear model (DI, fixed dt) ----
    STATE_KEYS = ("ro", "rou", "rov", "cxx", "cxy", "cyy")

    def _tangent(self, dU, dp=None):
        t = Tangent()
        assert lib().or_tangent_alloc(C.byref(t), self.d.np) == 0
        arrs = [np.ascontiguousarray(dU[k], dtype=np.float64) for k in self.STATE_KEYS]
        lib().or_tangent_set(C.byref(self.d), C.byref(self.st), C.byref(t),
                             *[_pd(a) for a in arrs])
        dp = dp or {}
        t.dWi, t.dbeta, t.deps = dp.get("Wi", 0.0), dp.get("beta", 0.0), dp.get("eps", 0.0)
        return t

    def rhs_tlm(self, dU, dp=None):
        """Tangent of calc_all_rhs (DI) at the current state along dU and the
        parameter direction dp = {"Wi", "beta", "eps"}."""
        t = self._tangent(dU, dp)
        n = self.d.npfb
        outs = [np.zeros(n) for _ in range(6)]
        ptrs = (PD * 6)(*[_pd(a) for a in outs])
        lib().or_calc_all_rhs_di_tlm(C.byref(self.d), C.byref(self.p), C.byref(self.st),
                                     C.byref(t), ptrs)
        lib().or_tangent_free(C.byref(t))
        return dict(zip(("ro", "rou", "rov", "xx", "xy", "yy"), outs))

    def rollout_tlm(self, dU, nsteps, dp=None):
        """Advance nsteps (fixed dt) with the tangent along dU and the parameter
        direction dp: (loss, dloss, final tangent)."""
        t = self._tangent(dU, dp)
        L, dL = C.c_double(), C.c_double()
        rc = lib().or_rollout_tlm(C.byref(self.d), C.byref(self.p), C.byref(self.st),
                                  C.byref(t), nsteps, C.byref(L), C.byref(dL))
        assert rc == 0
        n = self.d.npfb
        fin = {k: _view(getattr(t, k), n) for k in self.STATE_KEYS}
        lib().or_tangent_free(C.byref(t))
        return L.value, dL.value, fin

    def loss_trace(self):
        return lib().or_loss_trace(C.byref(self.d), C.byref(self.st))

    def diagnostics(self, kind):
        """statistics.F90 reductions: "mass_energy" (3), "conf" (6),
        "kolmogorov" (5), "tg" (1 value)."""
        d, p, st = C.byref(self.d), C.byref(self.p), C.byref(self.st)
        if kind == "mass_energy":
            out = np.zeros(3)
            lib().or_diag_mass_energy(d, p, st, _pd(out))
        elif kind == "conf":
            out = np.zeros(6)
            lib().or_diag_conf(d, st, _pd(out))
        elif kind == "kolmogorov":
            out = np.zeros(5)
            lib().or_diag_kolmogorov(d, p, st, _pd(out))
        else:
            out = np.array([lib().or_diag_tg(d, p, st)])
        return out

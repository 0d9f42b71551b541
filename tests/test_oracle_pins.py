"""Pins for the CPU oracle from closed forms, textbook routines and brute force.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails it.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as OL
from sunset_amd.nodes import jittered_lattice, taylor_green_ic


def _idx(order):
    return [(d - b, b) for d in range(1, order + 1) for b in range(d + 1)]


def test_wendland_closed_form_and_normalisation():
    L = OL.lib()
    h2 = 0.37 ** 2
    qs = np.linspace(-0.5, 2.5, 301)
    got = np.array([L.or_wab(q, h2) for q in qs])
    ref = np.where((qs > 0) & (qs <= 2), 7 / (4 * np.pi * h2) * (2 * qs + 1) * (1 - qs / 2) ** 4, 0)
    assert np.allclose(got, ref, rtol=1e-14, atol=0)
    # 2D normalisation: int_0^{2h} W 2 pi r dr = 1 when h^2 = h2 (Wendland C2)
    from scipy.integrate import quad
    h = 0.37
    val, _ = quad(lambda r: L.or_wab(r / h, h2) * 2 * np.pi * r, 0, 2 * h, limit=200)
    assert abs(val - 1.0) < 1e-10


@pytest.mark.parametrize("order", [4, 6, 8, 10])
def test_hermite_abfs_recurrence(order):
    """He_{n+1} = x He_n - n He_{n-1}; products He_p(x)He_q(y) in graded order."""
    rng = np.random.default_rng(1)
    n = (order * order + 3 * order) // 2
    out = np.zeros(n)
    for _ in range(20):
        x, y = rng.uniform(-2, 2, 2)
        OL.lib().or_abfs(order, x, y, out.ctypes.data_as(OL.PD))

        def he(k, z):
            a, b = 1.0, z
            if k == 0:
                return a
            for m in range(1, k):
                a, b = b, z * b - m * a
            return b
        ref = [he(a, x) * he(b, y) for a, b in _idx(order)]
        assert np.allclose(out, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("order", [4, 6, 8, 10])
def test_monomials_taylor(order):
    rng = np.random.default_rng(2)
    n = (order * order + 3 * order) // 2
    out = np.zeros(n)
    for _ in range(20):
        x, y = rng.uniform(-2, 2, 2)
        OL.lib().or_monomials(order, x, y, out.ctypes.data_as(OL.PD))
        ref = [x ** a * y ** b / (math.factorial(a) * math.factorial(b)) for a, b in _idx(order)]
        assert np.allclose(out, ref, rtol=1e-14, atol=1e-300)


def test_svd_solve_vs_lapack():
    rng = np.random.default_rng(3)
    for n in (5, 14, 27, 44):
        A = rng.standard_normal((n, n)) + n * np.eye(n) * 0.1
        b = rng.standard_normal(n)
        x = b.copy()
        Af = np.asfortranarray(A)
        rc = OL.lib().or_svd_solve(n, Af.ctypes.data_as(OL.PD), x.ctypes.data_as(OL.PD))
        assert rc == 0
        ref = np.linalg.solve(A, b)
        assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-11
    # exactly singular: zero singular value skipped -> minimum-norm solution
    A = np.diag([2.0, 1.0, 0.0])
    b = np.array([2.0, 3.0, 5.0])
    x = b.copy()
    OL.lib().or_svd_solve(3, np.asfortranarray(A).ctypes.data_as(OL.PD), x.ctypes.data_as(OL.PD))
    assert np.allclose(x, [1.0, 3.0, 0.0], atol=1e-15)


def test_eigens_and_log_conformation():
    rng = np.random.default_rng(4)
    L = OL.lib()
    l2 = (C.c_double * 2)()
    R = (C.c_double * 4)()
    for _ in range(200):
        M = rng.standard_normal((2, 2))
        S = M @ M.T + 0.1 * np.eye(2)
        a, b, d = S[0, 0], S[0, 1], S[1, 1]
        L.or_eigens2(a, b, d, l2, R)
        w = np.linalg.eigvalsh(S)
        assert np.allclose(sorted(l2), w, rtol=1e-12)
        Rm = np.array([[R[0], R[2]], [R[1], R[3]]])
        assert np.allclose(Rm @ np.diag(list(l2)) @ Rm.T, S, rtol=1e-12, atol=1e-13)
        assert l2[0] >= l2[1]
        # log-conformation round trip c -> psi -> c, and psi = logm(c)
        p = [C.c_double() for _ in range(3)]
        L.or_psi_from_c(a, b, d, *[C.byref(v) for v in p])
        ev, V = np.linalg.eigh(S)
        logS = V @ np.diag(np.log(ev)) @ V.T
        assert np.allclose([p[0].value, p[1].value, p[2].value],
                           [logS[0, 0], logS[0, 1], logS[1, 1]], atol=1e-12)
        c = [C.c_double() for _ in range(3)]
        L.or_c_from_psi(p[0].value, p[1].value, p[2].value, *[C.byref(v) for v in c])
        assert np.allclose([c[0].value, c[1].value, c[2].value], [a, b, d], rtol=1e-11)


def test_rk3_coefficients_order_conditions():
    """RK3(2)4[2R+]C (common_parameter.F90:75-94): the 2R low-storage form
    U_k = reg + a_k dt r_k, reg += b_k dt r_k is the Butcher table
    A[i][j] = b_j (j < i-1), A[i][i-1] = a_{i-1}.  Check order 3 for b and order
    2 for the embedded bhat."""
    a = [11847461282814 / 36547543011857, 3943225443063 / 7078155732230,
         -346793006927 / 4029903576067]
    b = np.array([1017324711453 / 9774461848756, 8237718856693 / 13685301971492,
                  57731312506979 / 19404895981398, -101169746363290 / 37734290219643])
    bh = np.array([15763415370699 / 46270243929542, 514528521746 / 5659431552419,
                   27030193851939 / 9429696342944, -69544964788955 / 30262026368149])
    A = np.zeros((4, 4))
    for i in range(1, 4):
        A[i, i - 1] = a[i - 1]
        A[i, : i - 1] = b[: i - 1]
    c = A.sum(1)
    assert abs(b.sum() - 1) < 1e-14 and abs(b @ c - 0.5) < 1e-14
    assert abs(b @ c ** 2 - 1 / 3) < 1e-14 and abs(b @ A @ c - 1 / 6) < 1e-14
    assert abs(bh.sum() - 1) < 1e-14 and abs(bh @ c - 0.5) < 1e-14
    assert abs(b @ c ** 2 - 1 / 3) < 1e-14
    # c1..c3 as common_parameter.F90:90-92; the reference sets c4 = 1 (:93) so
    # that `time` reads time0 + dt after the last stage (the RHS is autonomous)
    assert np.allclose(c[:3], [0, a[0], a[1] + b[0]], rtol=1e-14)


def _sim(order=4, n=24, hos=1.3, seed=7):
    ns = jittered_lattice(n, seed=seed, h_over_s=hos)
    return ns, OL.OracleSim(ns, order).setup()


def test_neighbours_vs_bruteforce():
    ns, o = _sim(order=6, n=20, hos=2.4)
    x, y = o.xy()
    cnt, cs, link = o.lists()
    for i in range(ns.npfb):
        dx = x[i] - x
        dy = y[i] - y
        r2 = dx * dx + dy * dy
        want = np.nonzero(r2 <= (ns.h[i] * 2.0) ** 2)[0]
        got = link[i, : cnt[i]]
        assert set(got.tolist()) == set(want.tolist())
        assert np.all(np.diff(r2[got]) >= 0)            # sorted by distance
        assert cs[i] == np.count_nonzero(r2 <= (ns.h_small[i] * 2.0) ** 2)


def test_periodic_images_cover_tiled_domain():
    """A ghost set is complete iff every node's stencil equals its stencil in a
    3x3-tiled copy of the domain."""
    ns, o = _sim(order=4, n=16, hos=1.3)
    x, y = o.xy()
    irel = o.irel()
    cnt, _, link = o.lists()
    tx = np.concatenate([ns.x + dx for dx in (-1, 0, 1) for _ in (-1, 0, 1)])
    ty = np.concatenate([ns.y + dy for _ in (-1, 0, 1) for dy in (-1, 0, 1)])
    for i in range(ns.npfb):
        r2 = (ns.x[i] - tx) ** 2 + (ns.y[i] - ty) ** 2
        want = sorted(np.nonzero(r2 <= (2 * ns.h[i]) ** 2 * (1 + 1e-12))[0] % ns.npfb)
        got = link[i, : cnt[i]]
        parents = sorted(np.where(got >= ns.npfb, irel[got], got))
        assert want == parents


@pytest.mark.parametrize("order,hos", [(4, 1.3), (6, 2.4), (8, 2.7)])
def test_weights_moment_conditions_and_polynomial_exactness(order, hos):
    ns, o = _sim(order=order, n=22, hos=hos)
    x, y = o.xy()
    cnt, cs, link = o.lists()
    W = o.weight_arrays()
    idx = _idx(order)
    worst = 0.0
    for i in range(0, ns.npfb, 5):
        js = link[i, : cnt[i]]
        X = (x[js] - x[i])
        Y = (y[js] - y[i])
        for m, (a, b) in enumerate(idx):
            mono = X ** a * Y ** b
            for comp, (da, db) in (("x", (1, 0)), ("y", (0, 1))):
                w = W["w_grad"][i, : cnt[i], 0 if comp == "x" else 1]
                exact = 1.0 if (a, b) == (da, db) else 0.0
                scale = np.abs(w).sum() * np.abs(mono).max()
                worst = max(worst, abs(w @ mono - exact) / max(scale, 1.0))
            lapx = 2.0 if (a, b) in ((2, 0), (0, 2)) else 0.0
            w = W["w_lap"][i, : cnt[i]]
            scale = np.abs(w).sum() * np.abs(mono).max()
            worst = max(worst, abs(w @ mono - lapx) / max(scale, 1.0))
        # sums
        assert np.allclose(W["gsum"][i], W["w_grad"][i, : cnt[i]].sum(0), rtol=1e-12, atol=1e-9)
    assert worst < 1e-9


def test_stencil_operators_are_linear_dense_identity():
    ns, o = _sim(order=4, n=16, hos=1.3)
    rng = np.random.default_rng(9)
    f = rng.standard_normal(ns.npfb)
    fe = o.extend(f)
    cnt, cs, link = o.lists()
    W = o.weight_arrays()
    irel = o.irel()
    n = ns.npfb
    D = np.zeros((n, n))
    for i in range(n):
        js = link[i, : cs[i]]
        par = np.where(js >= n, irel[js], js)
        np.add.at(D[i], par, W["w_grad"][i, : cs[i], 0])
        D[i, i] -= W["gsum"][i, 0]
    gx, gy = o.gradient(fe)
    assert np.allclose(gx, D @ f, rtol=1e-12, atol=1e-10)
    # filter term of a constant is zero; filtered constant is unchanged
    assert np.abs(o.filter_term(np.ones(o.np_))).max() < 1e-10
    assert np.allclose(o.filtered_var(np.full(o.np_, 3.0)), 3.0, rtol=0, atol=1e-12)


@pytest.mark.parametrize("order,expected", [(4, 4), (6, 6)])
def test_gradient_convergence_order(order, expected):
    errs = []
    for n in (16, 32):
        ns = jittered_lattice(n, seed=1, h_over_s={4: 1.8, 6: 2.4}[order])
        o = OL.OracleSim(ns, order).setup()
        x, y = o.xy()
        f = np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
        gx, _ = o.gradient(f)
        ex = 2 * np.pi * np.cos(2 * np.pi * ns.x) * np.cos(2 * np.pi * ns.y)
        errs.append(np.abs(gx - ex).max())
    rate = np.log2(errs[0] / errs[1])
    assert rate > expected - 1.0, (errs, rate)


def test_homogeneous_relaxation_closed_form_di_and_lc():
    """u = 0 and uniform c: dc/dt = -(1-2eps+eps tr c)(c-I)/lambda (DI, czz=1),
    and for LC d psi/dt = R diag(f/l_k - f) R^T / lambda."""
    ns, o = _sim(order=4, n=12, hos=1.3)
    n = ns.npfb
    c = np.array([[2.0, 0.3], [0.3, 1.5]])
    ic = dict(ro=np.ones(n), u=np.zeros(n), v=np.zeros(n), cxx=np.full(n, c[0, 0]),
              cxy=np.full(n, c[0, 1]), cyy=np.full(n, c[1, 1]))
    lam, eps = 0.7, 0.25
    o.make_phys(form="di", Wi=lam, eps=eps)
    o.set_initial(ic)
    r = o.rhs(1)
    fr = (1 - 2 * eps + eps * (c[0, 0] + c[1, 1] + 1.0)) / lam
    assert np.allclose(r["xx"], -fr * (c[0, 0] - 1), rtol=1e-9)
    assert np.allclose(r["xy"], -fr * c[0, 1], rtol=1e-9)
    assert np.allclose(r["yy"], -fr * (c[1, 1] - 1), rtol=1e-9)
    assert np.abs(r["ro"]).max() < 1e-9 and np.abs(r["rou"]).max() < 1e-7
    o2 = OL.OracleSim(ns, 4).setup()
    o2.make_phys(form="lc", Wi=lam, eps=eps)
    o2.set_initial(ic, itime=5)
    r2 = o2.rhs(2)
    ev, V = np.linalg.eigh(c)
    f = 1 - 2 * eps + eps * ev.sum()
    want = V @ np.diag(f / ev - f) @ V.T / lam
    assert np.allclose(r2["xx"], want[0, 0], rtol=1e-9)
    assert np.allclose(r2["xy"], want[0, 1], rtol=1e-9, atol=1e-12)
    assert np.allclose(r2["yy"], want[1, 1], rtol=1e-9)


def test_simple_shear_upper_convected_di():
    """Shear flow u = a sin(2 pi y), v = 0, c = I: the DI right-hand side reduces
    to the upper-convected term (grad u).c + c.(grad u)^T, so with the stencil's
    own velocity gradients rhs_xy = du/dy, rhs_xx = 2 du/dx, rhs_yy = 2 dv/dy."""
    ns = jittered_lattice(14, seed=3, h_over_s=1.3)
    o = OL.OracleSim(ns, 4).setup()
    n = ns.npfb
    a = 1e-3
    u = a * np.sin(2 * np.pi * ns.y)
    ic = dict(ro=np.ones(n), u=u, v=np.zeros(n), cxx=np.ones(n), cxy=np.zeros(n),
              cyy=np.ones(n))
    o.make_phys(form="di", Wi=1.0, eps=0.0)
    o.set_initial(ic)
    r = o.rhs(1)
    ux, uy = o.gradient(o.extend(u))
    assert np.allclose(r["xy"], uy, rtol=1e-12, atol=1e-18)
    assert np.allclose(r["xx"], 2 * ux, rtol=1e-12, atol=1e-18)
    assert np.abs(r["yy"]).max() == 0.0
    # and the stencil derivative is the analytic one to discretisation accuracy
    assert np.abs(uy - 2 * np.pi * a * np.cos(2 * np.pi * ns.y)).max() < 0.05 * 2 * np.pi * a


def test_lc_rhs_is_time_derivative_of_exp_psi():
    """Fattal-Kupferman: with the same discrete velocity gradient and a uniform
    conformation field, d/dt exp(psi) from the LC right-hand side equals the DI
    right-hand side (Oldroyd-B, eps = 0).  Checked by a central difference of
    c(psi) along rhs_psi."""
    ns = jittered_lattice(16, seed=11, h_over_s=1.8)
    n = ns.npfb
    u = 0.3 * np.sin(2 * np.pi * ns.x) * np.cos(2 * np.pi * ns.y)
    v = -0.2 * np.cos(2 * np.pi * ns.x) * np.sin(4 * np.pi * ns.y)
    c = np.array([[1.8, 0.4], [0.4, 1.2]])
    ic = dict(ro=np.ones(n), u=u, v=v, cxx=np.full(n, c[0, 0]), cxy=np.full(n, c[0, 1]),
              cyy=np.full(n, c[1, 1]))
    o_di = OL.OracleSim(ns, 4).setup()
    o_di.make_phys(form="di", Wi=0.8, eps=0.0)
    o_di.set_initial(ic)
    r_di = o_di.rhs(2)
    o_lc = OL.OracleSim(ns, 4).setup()
    o_lc.make_phys(form="lc", Wi=0.8, eps=0.0)
    o_lc.set_initial(ic, itime=4)
    r_lc = o_lc.rhs(2)
    F = o_lc.fields()
    L = OL.lib()
    d = 1e-6
    worst = 0.0
    for i in range(0, n, 7):
        out = []
        for sgn in (1, -1):
            cc = [C.c_double() for _ in range(3)]
            L.or_c_from_psi(F["pxx"][i] + sgn * d * r_lc["xx"][i],
                            F["pxy"][i] + sgn * d * r_lc["xy"][i],
                            F["pyy"][i] + sgn * d * r_lc["yy"][i], *[C.byref(q) for q in cc])
            out.append(np.array([q.value for q in cc]))
        dcdt = (out[0] - out[1]) / (2 * d)
        want = np.array([r_di["xx"][i], r_di["xy"][i], r_di["yy"][i]])
        worst = max(worst, np.abs(dcdt - want).max() / np.abs(want).max())
    assert worst < 1e-7, worst


def _moment_residuals(x, y, link, cnt, W, i, order):
    """max over monomials of degree <= order of |sum_k w_k X^a Y^b - exact| /
    (sum|w| max|X^a Y^b|), for the x-gradient and Laplacian weights of node i."""
    js = link[i, : cnt[i]]
    X, Y = x[js] - x[i], y[js] - y[i]
    out = 0.0
    for (a, b) in _idx(order):
        mono = X ** a * Y ** b
        for w, exact in ((W["w_grad"][i, : cnt[i], 0], 1.0 if (a, b) == (1, 0) else 0.0),
                         (W["w_lap"][i, : cnt[i]], 2.0 if (a, b) in ((2, 0), (0, 2)) else 0.0)):
            scale = np.abs(w).sum() * np.abs(mono).max()
            out = max(out, abs(w @ mono - exact) / max(scale, 1.0))
    return out


def test_mixed_order_weights_follow_the_order_map():
    """Cylinder channel (walls, no ghosts), order 8 inside and 4 near walls: each
    node's weights reproduce all monomials up to its own order, and the order-4
    nodes do not reproduce degree 5 (they were not solved at order 8)."""
    from sunset_amd.nodes import cylinder_channel
    ns, om = cylinder_channel(smin=0.04, smax=0.12, grow=1.0)
    assert set(np.unique(om)) == {4, 8}
    o = OL.OracleSim(ns, 8)
    o.ghosts()
    assert o.np_ == ns.npfb            # no periodic edges: no ghosts
    o.neighbours()
    o.weights_ordermap(om)
    x, y = o.xy()
    cnt, _, link = o.lists()
    W = o.weight_arrays()
    worst = max(_moment_residuals(x, y, link, cnt, W, i, int(om[i])) for i in range(ns.npfb))
    assert worst < 1e-9
    i4 = np.where(om == 4)[0][::7]
    deg5 = [abs(W["w_grad"][i, : cnt[i], 0] @ ((x[link[i, : cnt[i]]] - x[i]) ** 5))
            / (np.abs(W["w_grad"][i, : cnt[i], 0]).sum()
               * np.abs(x[link[i, : cnt[i]]] - x[i]).max() ** 5) for i in i4]
    assert np.median(deg5) > 1e-6


@pytest.mark.parametrize("order", [4, 6])
def test_convergence_order_equals_reference_vectors(order):
    """The oracle's operators reproduce the reference's errors on an analytic
    field at three resolutions (tests/golden/ref_conv_m*.npz), hence its
    observed convergence orders."""
    import os
    from parity_util import GOLDEN
    g = np.load(os.path.join(GOLDEN, f"ref_conv_m{order}.npz"))
    k = 2 * np.pi
    errs = []
    for n in g["sizes"]:
        ns = jittered_lattice(int(n), seed=int(g["seed"]), h_over_s=float(g["h_over_s"]))
        o = OL.OracleSim(ns, order).setup()
        x, y = o.xy()
        f = 1.0 + 0.1 * np.sin(k * x) * np.cos(k * y)
        gx, gy = o.gradient(f)
        lap = o.laplacian(f)
        ex = 0.1 * k * np.cos(k * ns.x) * np.cos(k * ns.y)
        ey = -0.1 * k * np.sin(k * ns.x) * np.sin(k * ns.y)
        el = -2 * k * k * 0.1 * np.sin(k * ns.x) * np.cos(k * ns.y)
        errs.append([np.abs(gx - ex).max(), np.abs(gy - ey).max(), np.abs(lap - el).max()])
    # the two sides' weights differ by ~1e-11 relative (SVD vs SVD), which
    # shows in the smallest errors as ~1e-10 absolute
    np.testing.assert_allclose(np.array(errs), g["errs"], rtol=1e-6, atol=1e-9)

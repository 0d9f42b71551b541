"""HIP path vs the reference (golden vectors from the flang build) and the CPU
oracle, through the C ABI.  Tolerances are in tests/parity_util.py:
neighbour lists and ghosts bit-exact; weights 1e-10 per node (max_k |dw| /
max_k |w|); operators 1e-12 on the same weights; fields after N RK3 steps 1e-9
relative to max(|f|, 1)."""
import numpy as np
import pytest

import oracle_lib as OL
from parity_util import (FIELD_RTOL, STENCIL_RTOL, WEIGHT_RTOL, csr_from_rows, field_rel,
                         golden_nodeset, golden_phys, ipart_nodeset, load_golden,
                         per_node_rel, taylor_green_ic)

pytestmark = pytest.mark.gpu

CASES = ["c1_m4", "m4_h18", "m6_20"]


@pytest.fixture(scope="module")
def sunset():
    import sunset_amd
    return sunset_amd


@pytest.fixture(scope="module", params=CASES)
def case(request, sunset):
    print(f"[gpu-parity] case {request.param}", flush=True)
    g = load_golden(request.param)
    ns = golden_nodeset(g)
    s = sunset.Solver(morder=int(g["order"]), form="lc")
    s.set_nodes(ns)
    cnt, cs, link = s.find_neighbours()
    s.calc_weights()
    s.calc_sums()
    s.filter_coefficients()
    return g, ns, s, (cnt, cs, link)


def test_ghosts_bit_exact(case):
    g, ns, s, _ = case
    x, y, ir, vr = s.ghosts()
    n = ns.npfb
    assert np.array_equal(x[n:], g["gx"]) and np.array_equal(y[n:], g["gy"])
    assert np.array_equal(ir, g["irel"])
    assert np.all(vr == 1)


def test_neighbour_lists_bit_exact(case):
    g, ns, s, (cnt, cs, link) = case
    x, y, _, _ = s.ghosts()
    n = ns.npfb
    assert np.array_equal(cnt, g["ij_count"]) and np.array_equal(cs, g["ij_count_small"])
    ptr, col, _ = csr_from_rows(cnt, link, x, y, x[:n], y[:n])
    assert np.array_equal(ptr, g["ptr"]) and np.array_equal(col, g["col"])
    # the device order is already the canonical (r^2, j) order
    for i in range(0, n, 37):
        assert np.array_equal(link[i, :cnt[i]], g["col"][g["ptr"][i]:g["ptr"][i + 1]])


def test_weights_match_reference(case):
    g, ns, s, (cnt, cs, link) = case
    x, y, _, _ = s.ghosts()
    n = ns.npfb
    W = s.weights(link.shape[1])
    ptr, col, (wx, wy, wl, wh) = csr_from_rows(
        cnt, link, x, y, x[:n], y[:n],
        [W["w_grad"][:, :, 0], W["w_grad"][:, :, 1], W["w_lap"], W["w_hyp"]])
    for a, b, name in ((wx, g["wx"], "x"), (wy, g["wy"], "y"), (wl, g["wl"], "lap"),
                       (wh, g["wh"], "hyp")):
        err = per_node_rel(ptr, a, b)
        print(f"weights {name}: max per-node rel {err:.3e}")
        assert err < WEIGHT_RTOL, name
    absx = np.add.reduceat(np.abs(g["wx"]), ptr[:-1])
    assert (np.abs(W["gsum"][:, 0] - g["gsum"][:, 0]) / absx).max() < WEIGHT_RTOL
    assert np.allclose(W["vol"], g["vol"], rtol=1e-13)


def test_weights_match_oracle(case):
    g, ns, s, (cnt, cs, link) = case
    o = OL.OracleSim(ns, int(g["order"])).setup()
    x, y, _, _ = s.ghosts()
    n = ns.npfb
    W = s.weights(link.shape[1])
    Wo = o.weight_arrays()
    oc, _, ol = o.lists()
    ptr, _, (wx, wh) = csr_from_rows(cnt, link, x, y, x[:n], y[:n],
                                     [W["w_grad"][:, :, 0], W["w_hyp"]])
    _, _, (ox, oh) = csr_from_rows(oc, ol, x, y, x[:n], y[:n],
                                   [Wo["w_grad"][:, :, 0], Wo["w_hyp"]])
    assert per_node_rel(ptr, wx, ox) < WEIGHT_RTOL
    assert per_node_rel(ptr, wh, oh) < WEIGHT_RTOL


def test_operators_exact_on_own_weights(case):
    """Stencil apply vs a float64 CSR matvec on the device's own weights."""
    g, ns, s, (cnt, cs, link) = case
    rng = np.random.default_rng(4)
    x, y, ir, _ = s.ghosts()
    n = ns.npfb
    f = np.sin(2 * np.pi * x) * np.cos(4 * np.pi * y) + 0.1 * rng.standard_normal(s.np)
    W = s.weights(link.shape[1])
    gx, gy = s.gradient(f)
    lap = s.laplacian(f)
    ft = s.filter_term(f)
    fv = s.filtered_var(f)
    g1 = s.gradient_only(1, f)
    g2 = s.gradient_only(2, f)
    ref_gx = np.array([f[link[i, :cs[i]]] @ W["w_grad"][i, :cs[i], 0] - f[i] * W["gsum"][i, 0]
                       for i in range(n)])
    ref_gy = np.array([f[link[i, :cs[i]]] @ W["w_grad"][i, :cs[i], 1] - f[i] * W["gsum"][i, 1]
                       for i in range(n)])
    ref_l = np.array([f[link[i, :cs[i]]] @ W["w_lap"][i, :cs[i]] - f[i] * W["lsum"][i]
                      for i in range(n)])
    ref_h = np.array([f[link[i, :cnt[i]]] @ W["w_hyp"][i, :cnt[i]] - f[i] * W["hsum"][i]
                      for i in range(n)])
    sc = lambda w, c: np.array([np.abs(w[i, :c[i]]).sum() for i in range(n)]) * np.abs(f).max()
    for got, ref, scale in ((gx, ref_gx, sc(W["w_grad"][:, :, 0], cs)),
                            (gy, ref_gy, sc(W["w_grad"][:, :, 1], cs)),
                            (g1, ref_gx, sc(W["w_grad"][:, :, 0], cs)),
                            (g2, ref_gy, sc(W["w_grad"][:, :, 1], cs)),
                            (lap, ref_l, sc(W["w_lap"], cs)),
                            (ft, ref_h, sc(W["w_hyp"], cnt))):
        assert (np.abs(got - ref) / scale).max() < STENCIL_RTOL
    assert np.allclose(fv, f[:n] + ref_h, rtol=0, atol=1e-12 * np.abs(f).max())
    assert np.array_equal(s.gradient_only(3, f), np.zeros(n))


def test_rhs_matches_oracle(case):
    g, ns, s, _ = case
    ph = golden_phys(g)
    ic = taylor_green_ic(ns)
    ic["cxx"] = 1.0 + 0.3 * np.cos(2 * np.pi * ns.x) ** 2
    ic["cxy"] = 0.2 * np.sin(2 * np.pi * ns.y)
    for form in ("lc", "di"):
        for itime, rk in ((3, 1), (1, 1), (1, 2)):
            import sunset_amd
            d = sunset_amd.Solver(morder=int(g["order"]), form=form).setup(ns)
            d.set_phys(**ph, Mdiff=1e-3 if form == "di" else 0.0)
            d.set_initial(ic, itime=itime)
            r = d.rhs(rk)
            o = OL.OracleSim(ns, int(g["order"])).setup()
            o.make_phys(form=form, **ph, Mdiff=1e-3 if form == "di" else 0.0)
            o.set_initial(ic, itime=itime)
            ro = o.rhs(rk)
            for k in r:
                scale = max(np.abs(ro[k]).max(), 1.0)
                assert np.abs(r[k] - ro[k]).max() / scale < 1e-9, (form, itime, rk, k)
            d.close()


def test_fields_after_rk3_steps_match_reference(case):
    g, ns, s, _ = case
    s.set_phys(**golden_phys(g))
    s.set_initial(taylor_green_ic(ns))
    s.run(int(g["nsteps"]))
    F = s.fields()
    assert F["itime"] == int(g["itime"])
    worst = 0.0
    for k in ("ro", "rou", "rov", "u", "v", "pxx", "pxy", "pyy", "cxx", "cxy", "cyy"):
        e = field_rel(F[k], g[f"f_{k}"])
        worst = max(worst, e)
        assert e < FIELD_RTOL, k
    ctl = g["ctl"]
    assert abs(F["dt"] - ctl[0]) <= 1e-10 * ctl[0]
    assert abs(F["time"] - ctl[1]) <= 1e-10 * ctl[1]
    assert abs(F["emax_np1"] - ctl[2]) <= 1e-3 * ctl[2] + 1e-15
    print(f"fields after {int(g['nsteps'])} steps: max rel {worst:.3e}")


def test_di_fields_match_oracle_and_graph_replay_matches_eager(sunset):
    ns = golden_nodeset(load_golden("m4_h18"))
    ic = taylor_green_ic(ns)
    ic["cxx"] = 1.0 + 0.2 * np.cos(2 * np.pi * ns.x) ** 2
    ph = dict(Re=1.0, Wi=1.0, beta=0.5, eps=0.0, Ma=0.1, Mdiff=1e-3)
    out = {}
    for graph in (False, True):
        s = sunset.Solver(morder=4, form="di").setup(ns)
        s.set_phys(**ph)
        s.set_initial(ic)
        for rep in range(3):       # replays separated by host synchronisation
            s.run(10, graph=graph)
            _ = s.fields()
        out[graph] = s.fields()
        s.close()
    o = OL.OracleSim(ns, 4).setup()
    o.make_phys(form="di", **ph)
    o.set_initial(ic)
    o.run(30)
    R = o.fields()
    for k in ("ro", "rou", "rov", "cxx", "cxy", "cyy"):
        assert field_rel(out[False][k], R[k]) < FIELD_RTOL, k
        assert np.array_equal(out[False][k], out[True][k]), k
    assert out[True]["itime"] == 30


def test_ipart_fixture_lists_and_weights(sunset):
    ns = ipart_nodeset(order=4)
    s = sunset.Solver(morder=4, form="lc")
    s.set_nodes(ns)
    cnt, cs, link = s.find_neighbours()
    s.calc_weights()
    s.calc_sums()
    s.filter_coefficients()
    o = OL.OracleSim(ns, 4).setup()
    oc, ocs, ol = o.lists()
    assert np.array_equal(cnt, oc) and np.array_equal(cs, ocs)
    for i in range(ns.npfb):
        assert np.array_equal(link[i, :cnt[i]], ol[i, :cnt[i]])
    W = s.weights(link.shape[1])
    Wo = o.weight_arrays()
    ptr = np.concatenate([[0], np.cumsum(cnt)])
    pack = lambda w, c: np.concatenate([w[i, :c[i]] for i in range(len(c))])
    assert per_node_rel(ptr, pack(W["w_lap"], cnt), pack(Wo["w_lap"], oc)) < WEIGHT_RTOL


def test_loop_order_does_not_change_results(sunset):
    """The time loop runs its nodes in Morton order (cache locality); the
    caller's order gives the same fields up to the summation order of the mass
    correction (one reduction per step)."""
    g = load_golden("m6_20")
    ns = golden_nodeset(g)
    out = []
    for order in (0, 1):
        s = sunset.Solver(morder=6, form="lc")
        assert lib_set_option(s, "loop_order", order) == 0
        s.setup(ns)
        s.set_phys(**golden_phys(g))
        s.set_initial(taylor_green_ic(ns))
        s.run(10)
        out.append(s.fields())
        s.close()
    for k in ("ro", "rou", "rov", "pxx", "pxy", "pyy", "cxx"):
        assert field_rel(out[0][k], out[1][k]) < 1e-12, k
    assert out[0]["dt"] == out[1]["dt"]


def lib_set_option(s, key, value):
    import sunset_amd
    return sunset_amd.lib().sv_set_option(s.h, key.encode(), float(value))


def test_halo_sweeps_match_global_gathers_bitwise(sunset):
    """The LDS halo-staged sweeps read the same records in the same order as
    the global-gather sweeps: identical fields, LC and DI."""
    g = load_golden("m6_20")
    ns = golden_nodeset(g)
    for form in ("lc", "di"):
        out = []
        for halo in (0, 1):
            s = sunset.Solver(morder=6, form=form)
            assert lib_set_option(s, "halo_sweeps", halo) == 0
            s.setup(ns)
            s.set_phys(**golden_phys(g), Mdiff=1e-3 if form == "di" else 0.0)
            s.set_initial(taylor_green_ic(ns))
            s.run(6)
            out.append(s.fields())
            s.close()
        for k in ("ro", "rou", "rov", "cxx", "cxy", "cyy"):
            assert np.array_equal(out[0][k], out[1][k]), (form, k)

"""Boundary nodes on the device against the oracle: the boundary operators of
calc_boundary_weights, the drop-in derivative operators on a node set with
walls / inflow / outflow, the right-hand side with characteristic boundary
conditions, and RK3 steps (PID time steps, eager and captured) in the
log-conformation form.  The oracle's boundary restatement is pinned against the
compiled reference in test_oracle_bound.py.

Tolerances: operators of boundary nodes and rows 1-2 (6x6 systems, LU here,
SVD in the oracle) to 1e-9 of the row's largest weight; LABFM weights of
interior nodes as the periodic-lattice weight tests (1e-9), except order 8
within a stencil radius of a wall, whose one-sided moment matrices reach
cond ~1e9 (1e-6); fields after 20 steps to 1e-9 of max(|field|, 1e-2)."""
import numpy as np
import pytest

import os

import oracle_lib as OL
import refdrv
from parity_util import FIELD_RTOL, field_rel, golden_bcase
from sunset_amd.nodes import inflow_outflow_channel, wall_channel

pytestmark = pytest.mark.gpu

PH = dict(Re=10.0, Wi=0.5, beta=0.5, eps=0.1, Ma=0.1)


def _sv():
    import sunset_amd
    return sunset_amd


def wall_ic(ns):
    n = ns.npfb
    return dict(ro=np.ones(n), u=4 * ns.y * (1 - ns.y),
                v=0.01 * np.sin(2 * np.pi * ns.x) * np.sin(np.pi * ns.y) ** 2,
                cxx=np.ones(n), cxy=np.zeros(n), cyy=np.ones(n))


def io_ic(ns):
    n = ns.npfb
    return dict(ro=np.ones(n), u=1.0 + 0.01 * np.sin(2 * np.pi * ns.y),
                v=0.01 * np.cos(2 * np.pi * ns.y) * np.sin(np.pi * ns.x),
                cxx=np.ones(n), cxy=np.zeros(n), cyy=np.ones(n))


def _node_set(kind, order, n=20):
    return (wall_channel(n, n, order=order, seed=1) if kind == "wall"
            else inflow_outflow_channel(n, n, order=order, seed=2))


@pytest.mark.parametrize("kind,order", [("wall", 4), ("wall", 6), ("wall", 8), ("io", 6)])
def test_boundary_weights_vs_oracle(kind, order):
    sv = _sv()
    ns = _node_set(kind, order, 16)
    print(f"[bound] weights {kind} order {order} npfb {ns.npfb}", flush=True)
    s = sv.Solver(morder=order, form="lc").setup(ns)
    o = OL.OracleSim(ns, order).setup()
    cnt_o, _, link_o = o.lists()
    Wo = o.weight_arrays()
    # nplink is the longest list before the boundary pruning (labf.F90:69)
    Wd = s.weights(2 * int(cnt_o.max()))
    s.close()
    worst = 0.0
    for i in range(ns.npfb):
        t = int(ns.node_type[i])
        tol = 1e-9 if (0 <= t <= 2 or t in (-1, -2) or order < 8) else 1e-6
        K = cnt_o[i]
        for key, comp in (("w_grad", 0), ("w_grad", 1), ("w_lap", None), ("w_hyp", None)):
            a = Wd[key][i, :K] if comp is None else Wd[key][i, :K, comp]
            b = Wo[key][i, :K] if comp is None else Wo[key][i, :K, comp]
            # both lists are sorted by (r^2, j), so entries correspond
            sc = max(np.abs(b).max(), 1e-300)
            err = np.abs(a - b).max() / sc
            worst = max(worst, err)
            assert err < tol, (i, t, key, comp, err)
        assert np.all(Wd["w_lap"][i, K:] == 0.0) and np.all(Wd["w_hyp"][i, K:] == 0.0), i
    for k in ("gsum", "lsum", "hsum", "vol"):
        err = np.abs(Wd[k] - Wo[k]).max() / np.abs(Wo[k]).max()
        assert err < (1e-9 if order < 8 else 1e-6), (k, err)
    print(f"[bound] worst row-relative difference {worst:.2e}", flush=True)


def test_boundary_lists_pruned():
    """A boundary node keeps its same-type neighbours and its FD rows
    (labf.F90:541-566); every other row is the neighbour search's."""
    sv = _sv()
    ns = _node_set("wall", 6, 16)
    s = sv.Solver(morder=6, form="lc")
    s.set_nodes(ns)
    cnt0, cs0, link0 = s.find_neighbours()
    s.calc_weights()
    s.calc_boundary_weights()
    s.calc_sums()
    s.filter_coefficients()
    o = OL.OracleSim(ns, 6).setup()
    cnt_o, cs_o, link_o = o.lists()
    x, y, irel, vrel = s.ghosts()
    nt_all = np.concatenate([ns.node_type, ns.node_type[irel]])
    for i in np.where((ns.node_type >= 0) & (ns.node_type <= 2))[0]:
        full = link0[i, :cnt0[i]]
        keep = [j for j in full if nt_all[j] == ns.node_type[i] or
                (j < ns.npfb and ns.node_type[j] < 0 and j + ns.node_type[j] == i)]
        assert list(keep) == list(link_o[i, :cnt_o[i]]), i
    s.close()


@pytest.mark.parametrize("kind", ["wall", "io"])
def test_boundary_operators_vs_oracle(kind):
    """Drop-in calc_gradient / calc_laplacian / calc_filtered_var on a node set
    with boundary nodes (derivatives.F90:26-116, :360-414)."""
    sv = _sv()
    ns = _node_set(kind, 6, 16)
    s = sv.Solver(morder=6, form="lc").setup(ns)
    o = OL.OracleSim(ns, 6).setup()
    x, y = o.xy()
    phi = np.sin(2 * np.pi * x) * np.cos(np.pi * y) + y ** 3
    gx, gy = s.gradient(phi)
    ox, oy = o.gradient(phi)
    for a, b in ((gx, ox), (gy, oy), (s.laplacian(phi), o.laplacian(phi)),
                 (s.laplacian_transverse(phi), o.laplacian_transverse(phi)),
                 (s.filtered_var(phi), o.filtered_var(phi))):
        assert np.abs(a - b).max() <= 1e-9 * max(np.abs(b).max(), 1.0)
    # calc_laplacian_transverse_only_on_bound differs from the Laplacian on
    # the boundary nodes only
    bnd = (ns.node_type >= 0) & (ns.node_type <= 2)
    lt, lap = s.laplacian_transverse(phi), s.laplacian(phi)
    assert np.array_equal(lt[~bnd], lap[~bnd])
    assert np.abs(lt[bnd] - lap[bnd]).max() > 1e-3 * np.abs(lap[bnd]).max()
    s.close()


def _run_pair(kind, order, inflow, nsteps, graph=False):
    sv = _sv()
    ns = _node_set(kind, order)
    ic = wall_ic(ns) if kind == "wall" else io_ic(ns)
    drive = (8.0, 0.0) if kind == "wall" else (0.0, 0.0)
    s = sv.Solver(morder=order, form="lc").setup(ns)
    s.set_inflow(**inflow)
    s.set_phys(drive=(*drive, 0.0), **PH)
    s.set_initial(ic)
    s.run(nsteps, graph=graph)
    F = s.fields()
    s.close()
    o = OL.OracleSim(ns, order).setup()
    o.make_phys(form="lc", drive=drive, **PH, **inflow)
    o.set_initial(ic)
    o.run(nsteps)
    return ns, F, o.fields()


@pytest.mark.parametrize("kind,order,inflow", [
    ("wall", 6, {}), ("wall", 4, {}),
    ("io", 6, dict(inflow_type=0)), ("io", 6, dict(inflow_type=1)),
    ("io", 6, dict(inflow_type=2)),
    ("io", 4, dict(inflow_type=1, u_control=1, u_start=0.5, u_end=1.0, ramp_time=1e-4))])
def test_boundary_stepping_vs_oracle(kind, order, inflow):
    print(f"[bound] steps {kind} order {order} {inflow}", flush=True)
    ns, F, R = _run_pair(kind, order, inflow, 20)
    assert F["dt"] == pytest.approx(R["dt"], rel=1e-12)
    for k in ("ro", "rou", "rov", "pxx", "pxy", "pyy"):
        err = np.abs(F[k] - R[k]).max() / max(np.abs(R[k]).max(), 1e-2)
        assert err < 1e-9, (k, err)
    walls = ns.node_type == 0
    assert np.all(F["rou"][walls] == 0.0) and np.all(F["rov"][walls] == 0.0)


def test_boundary_graph_equals_eager():
    ns, F, _ = _run_pair("io", 6, dict(inflow_type=1, u_control=1, u_start=0.5, u_end=1.0,
                                       ramp_time=1e-4), 12)
    _, G, _ = _run_pair("io", 6, dict(inflow_type=1, u_control=1, u_start=0.5, u_end=1.0,
                                      ramp_time=1e-4), 12, graph=True)
    for k in ("ro", "rou", "rov", "pxx", "pxy", "pyy"):
        assert np.array_equal(F[k], G[k]), k


@pytest.mark.parametrize("kind", ["wall", "io"])
def test_boundary_rhs_vs_oracle(kind):
    """calc_all_rhs on the initial state, first stage of the first step
    (the rhs_lc.F90:503-506 branch) and a later stage."""
    sv = _sv()
    ns = _node_set(kind, 6, 16)
    ic = wall_ic(ns) if kind == "wall" else io_ic(ns)
    s = sv.Solver(morder=6, form="lc").setup(ns)
    s.set_phys(**PH)
    s.set_initial(ic, itime=1)
    o = OL.OracleSim(ns, 6).setup()
    o.make_phys(form="lc", **PH)
    o.set_initial(ic, itime=1)
    for stage in (1, 2):
        a = s.rhs(stage)
        b = o.rhs(stage)
        for k in a:
            err = np.abs(a[k] - b[k]).max() / max(np.abs(b[k]).max(), 1e-2)
            assert err < 1e-9, (stage, k, err)
    s.close()


def test_boundary_errors():
    sv = _sv()
    ns = _node_set("wall", 6, 16)
    # the adjoint takes no boundary nodes
    s = sv.Solver(morder=6, form="di").setup(ns)
    s.set_phys(**PH)
    s.set_initial(wall_ic(ns), dt=1e-6, dt_mode="fixed")
    with pytest.raises(sv.SunsetError, match="boundary"):
        s.adj_setup(5)
    s.close()
    # sums before the boundary operators
    s = sv.Solver(morder=6, form="lc")
    s.set_nodes(ns)
    s.find_neighbours(copy=False)
    s.calc_weights()
    with pytest.raises(sv.SunsetError, match="sv_calc_boundary_weights"):
        s.calc_sums()
    s.close()
    # an FD row out of place
    bad = _node_set("wall", 6, 16)
    bad.node_type = bad.node_type.copy()
    bad.node_type[2] = -3
    s = sv.Solver(morder=6, form="lc")
    with pytest.raises(sv.SunsetError, match="FD row"):
        s.set_nodes(bad)
    s.close()
    # a non-unit normal
    bad = _node_set("wall", 6, 16)
    bad.rny = bad.rny * 2.0
    s = sv.Solver(morder=6, form="lc")
    with pytest.raises(sv.SunsetError, match="unit vector"):
        s.set_nodes(bad)
    s.close()


@pytest.mark.parametrize("name", ["wall_m6", "io_m6_soft", "io_m4_ramp"])
def test_boundary_device_vs_golden(name):
    """The committed vectors of the compiled reference: weights (CSR, canonical
    order) and the fields after 40 RK3 steps, PID time steps.  psi tolerance
    as in test_oracle_bound.test_boundary_oracle_vs_golden (near-isotropic
    outflow nodes)."""
    sv = _sv()
    g, ns, ic, phys, drive, inflow = golden_bcase(name)
    order = int(g["order"])
    print(f"[bound] golden {name}", flush=True)
    s = sv.Solver(morder=order, form="lc").setup(ns)
    W = s.weights(2 * int(g["ij_count"].max()))
    ptr = g["ptr"]
    for i in range(ns.npfb):
        K = int(g["ij_count"][i])
        sl = slice(ptr[i], ptr[i + 1])
        for a, gk in ((W["w_grad"][i, :K, 0], "wx"), (W["w_grad"][i, :K, 1], "wy"),
                      (W["w_lap"][i, :K], "wl"), (W["w_hyp"][i, :K], "wh")):
            b = g[gk][sl]
            assert np.abs(a - b).max() <= 1e-9 * max(np.abs(b).max(), 1e-300), (i, gk)
    s.set_inflow(**inflow)
    s.set_phys(drive=(*drive, 0.0), **phys)
    s.set_initial(ic)
    s.run(int(g["nsteps"]), graph=True)
    F = s.fields()
    s.close()
    assert F["dt"] == pytest.approx(float(g["ctl"][0]), rel=1e-12)
    for k in ("ro", "rou", "rov"):
        assert field_rel(F[k], g[f"f_{k}"]) < FIELD_RTOL, k
    for k in ("pxx", "pxy", "pyy"):
        assert field_rel(F[k], g[f"f_{k}"]) < 1e-8, k


@pytest.mark.parametrize("name", ["wall_m6", "io_m6_soft", "io_m4_ramp"])
def test_boundary_dropin_vs_golden(name):
    """The reference's own time loop with this library's Fortran drop-in
    neighbours / labf (calc_boundary_weights included) / derivatives modules."""
    g, ns, ic, phys, drive, inflow = golden_bcase(name)
    order = int(g["order"])
    binary = refdrv.dropin_binary(order)
    if not os.path.exists(binary):
        pytest.skip("drop-in binary not built (make -C oracle dropin)")
    print(f"[bound] drop-in {name}", flush=True)
    out, _ = refdrv.run_reference(ns, order, ic, nsteps=int(g["nsteps"]),
                                  phys=dict(phys, drive=(*drive, 0.0)), inflow=inflow,
                                  binary=binary, timeout=300, fields_only=True)
    for k in ("ro", "rou", "rov"):
        assert field_rel(out[k], g[f"f_{k}"]) < FIELD_RTOL, k
    for k in ("pxx", "pxy", "pyy"):
        assert field_rel(out[k], g[f"f_{k}"]) < 1e-8, k


def test_curved_wall_vs_oracle():
    """A no-slip cylinder in a periodic box, order 6 (curvature terms, wall
    frame rotation): weights and 20 RK3 steps, device against the oracle
    (the oracle is checked against the compiled reference on this case in
    test_oracle_bound.test_curved_wall_vs_reference)."""
    sv = _sv()
    from sunset_amd.nodes import cylinder_in_box
    ns = cylinder_in_box(32, order=6, seed=3)
    n = ns.npfb
    ic = dict(ro=np.ones(n), u=np.where(ns.node_type == 0, 0.0, 0.1 * np.sin(2 * np.pi * ns.y)),
              v=np.zeros(n), cxx=np.ones(n), cxy=np.zeros(n), cyy=np.ones(n))
    print(f"[bound] cylinder in a box, {n} nodes", flush=True)
    s = sv.Solver(morder=6, form="lc").setup(ns)
    o = OL.OracleSim(ns, 6).setup()
    cnt, _, _ = o.lists()
    Wo = o.weight_arrays()
    Wd = s.weights(2 * int(cnt.max()))
    for i in range(n):
        K = cnt[i]
        for key in ("w_grad", "w_lap", "w_hyp"):
            b = Wo[key][i, :K]
            assert np.abs(Wd[key][i, :K] - b).max() <= 1e-9 * max(np.abs(b).max(), 1e-300), (i, key)
    s.set_phys(drive=(1.0, 0.0, 0.0), **PH)
    s.set_initial(ic)
    s.run(20)
    F = s.fields()
    s.close()
    o.make_phys(form="lc", drive=(1.0, 0.0), **PH)
    o.set_initial(ic)
    o.run(20)
    R = o.fields()
    assert F["dt"] == pytest.approx(R["dt"], rel=1e-12)
    for k in ("ro", "rou", "rov", "pxx", "pxy", "pyy"):
        assert field_rel(F[k], R[k]) < FIELD_RTOL, k


@pytest.mark.parametrize("kind,inflow", [("wall", {}), ("io", dict(inflow_type=0)),
                                         ("io", dict(inflow_type=1))])
def test_boundary_direct_form_vs_oracle(kind, inflow):
    """The direct (DI) form with boundary nodes: walls with the rotated
    upper-convected terms, inflow/outflow conformation right-hand sides zero as
    the reference's DI branch (rhs_di.F90:477-594); 20 RK3 steps against the
    oracle (whose DI boundary branch is pinned to the reference-pinned LC one
    in test_oracle_bound.test_di_boundary_rhs_pinned_by_lc)."""
    sv = _sv()
    ns = _node_set(kind, 6)
    ic = wall_ic(ns) if kind == "wall" else io_ic(ns)
    drive = (8.0, 0.0) if kind == "wall" else (0.0, 0.0)
    print(f"[bound] DI {kind} {inflow}", flush=True)
    s = sv.Solver(morder=6, form="di").setup(ns)
    s.set_inflow(**inflow)
    s.set_phys(drive=(*drive, 0.0), **PH)
    s.set_initial(ic)
    s.run(20)
    F = s.fields()
    s.close()
    o = OL.OracleSim(ns, 6).setup()
    o.make_phys(form="di", drive=drive, **PH, **inflow)
    o.set_initial(ic)
    o.run(20)
    R = o.fields()
    assert F["dt"] == pytest.approx(R["dt"], rel=1e-12)
    for k in ("ro", "rou", "rov", "cxx", "cxy", "cyy"):
        assert field_rel(F[k], R[k]) < FIELD_RTOL, k


def _varying_c_ic(ns, kind):
    ic = wall_ic(ns) if kind == "wall" else io_ic(ns)
    t = np.minimum(ns.y, 1.0 - ns.y) if kind == "wall" else ns.x
    ic["cxx"] = 1.0 + 0.3 * t + 0.2 * np.sin(2 * np.pi * ns.x) * t ** 2
    ic["cxy"] = 0.1 * t - 0.05 * t ** 3
    ic["cyy"] = 1.0 + 0.2 * t ** 2
    return ic


@pytest.mark.parametrize("kind", ["wall", "io"])
def test_boundary_direct_form_mdiff_rhs_vs_oracle(kind):
    """Direct form with polymer diffusion: on walls the wall Laplacian of c
    plus the one-sided no-flux term over the FD rows (rhs_di.F90:528-556); the
    right-hand side of a conformation field varying along the normal, device
    against the oracle (whose term is pinned by a closed form in
    test_oracle_bound.test_di_wall_polymer_diffusion_closed_form)."""
    sv = _sv()
    ns = _node_set(kind, 6)
    ic = _varying_c_ic(ns, kind)
    print(f"[bound] DI Mdiff rhs {kind}", flush=True)
    s = sv.Solver(morder=6, form="di").setup(ns)
    s.set_phys(Mdiff=2e-3, **PH)
    s.set_initial(ic)
    o = OL.OracleSim(ns, 6).setup()
    o.make_phys(form="di", Mdiff=2e-3, **PH)
    o.set_initial(ic)
    a = s.rhs(1)
    b = o.rhs(1)
    walls = ns.node_type == 0
    for k in a:
        err = np.abs(a[k] - b[k]).max() / max(np.abs(b[k]).max(), 1e-2)
        assert err < 1e-9, (k, err)
    if kind == "wall":
        # the wall term is present: the walls' conformation rhs moves with Mdiff
        s0 = sv.Solver(morder=6, form="di").setup(ns)
        s0.set_phys(**PH)
        s0.set_initial(ic)
        assert np.abs(s0.rhs(1)["xx"][walls] - a["xx"][walls]).max() > 1e-3
        s0.close()
    s.close()


@pytest.mark.parametrize("kind,inflow", [("wall", {}), ("io", dict(inflow_type=0))])
def test_boundary_direct_form_mdiff_stepping_vs_oracle(kind, inflow):
    sv = _sv()
    ns = _node_set(kind, 6)
    ic = _varying_c_ic(ns, kind)
    drive = (8.0, 0.0) if kind == "wall" else (0.0, 0.0)
    print(f"[bound] DI Mdiff steps {kind} {inflow}", flush=True)
    s = sv.Solver(morder=6, form="di").setup(ns)
    s.set_inflow(**inflow)
    s.set_phys(drive=(*drive, 0.0), Mdiff=2e-3, **PH)
    s.set_initial(ic)
    s.run(20)
    F = s.fields()
    s.close()
    o = OL.OracleSim(ns, 6).setup()
    o.make_phys(form="di", drive=drive, Mdiff=2e-3, **PH, **inflow)
    o.set_initial(ic)
    o.run(20)
    R = o.fields()
    assert F["dt"] == pytest.approx(R["dt"], rel=1e-12)
    for k in ("ro", "rou", "rov", "cxx", "cxy", "cyy"):
        assert field_rel(F[k], R[k]) < FIELD_RTOL, k


@pytest.mark.parametrize("kind,form,mdiff", [("wall", "lc", 0.0), ("wall", "di", 2e-3),
                                             ("io", "lc", 0.0), ("io", "di", 2e-3)])
def test_boundary_halo_sweeps_match_global_gathers(kind, form, mdiff):
    """The boundary variants of the stage sweep with LDS halos and with global
    gathers (halo_sweeps 0) give identical fields."""
    sv = _sv()
    ns = _node_set(kind, 6)
    ic = _varying_c_ic(ns, kind)
    drive = (8.0, 0.0, 0.0) if kind == "wall" else (0.0, 0.0, 0.0)
    print(f"[bound] halo vs global {kind} {form} Mdiff {mdiff}", flush=True)
    out = []
    for halo in (0, 1):
        s = sv.Solver(morder=6, form=form)
        assert sv.lib().sv_set_option(s.h, b"halo_sweeps", float(halo)) == 0
        s.setup(ns)
        if kind == "io":
            s.set_inflow(inflow_type=1, u_control=1, u_start=0.5, u_end=1.0, ramp_time=1e-4)
        s.set_phys(drive=drive, Mdiff=mdiff, **PH)
        s.set_initial(ic)
        s.run(8)
        out.append(s.fields())
        s.close()
    keys = ("ro", "rou", "rov") + (("pxx", "pxy", "pyy") if form == "lc" else ("cxx", "cxy", "cyy"))
    for k in keys:
        assert np.array_equal(out[0][k], out[1][k]), k


def test_curved_wall_direct_form_mdiff_rhs_vs_oracle():
    """The direct form's wall diffusion on a curved wall (the normal part of
    the wall Laplacian carries the curvature term, labf.F90:1011): the
    right-hand side of a conformation field varying in x and y, device against
    the oracle."""
    sv = _sv()
    from sunset_amd.nodes import cylinder_in_box
    ns = cylinder_in_box(32, order=6, seed=3)
    n = ns.npfb
    x, y = ns.x, ns.y
    ic = dict(ro=np.ones(n), u=np.where(ns.node_type == 0, 0.0, 0.1 * np.sin(2 * np.pi * y)),
              v=np.zeros(n), cxx=1.0 + 0.2 * np.cos(2 * np.pi * x) * np.sin(2 * np.pi * y) ** 2,
              cxy=0.1 * np.sin(2 * np.pi * (x + y)), cyy=1.0 + 0.1 * np.cos(2 * np.pi * y))
    print(f"[bound] cylinder DI Mdiff rhs, {n} nodes", flush=True)
    s = sv.Solver(morder=6, form="di").setup(ns)
    s.set_phys(Mdiff=2e-3, **PH)
    s.set_initial(ic)
    o = OL.OracleSim(ns, 6).setup()
    o.make_phys(form="di", Mdiff=2e-3, **PH)
    o.set_initial(ic)
    a = s.rhs(1)
    b = o.rhs(1)
    s.close()
    for k in a:
        err = np.abs(a[k] - b[k]).max() / max(np.abs(b[k]).max(), 1e-2)
        assert err < 1e-9, (k, err)

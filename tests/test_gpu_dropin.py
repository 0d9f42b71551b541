"""The Fortran drop-in: the reference's own RK3 loop (rhs_lc, step, mirrors),
built with this library's drop-in `neighbours`, `labf` and `derivatives`
modules (iso_c_binding over the C ABI) in place of the reference's, against
the unmodified reference build and the committed golden vectors."""
import os

import numpy as np
import pytest

import refdrv
from parity_util import (FIELD_RTOL, WEIGHT_RTOL, csr_from_rows, field_rel, golden_nodeset,
                         golden_phys, load_golden, per_node_rel, taylor_green_ic)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["m4_h18", "m6_20"])
def test_dropin_matches_golden(name):
    g = load_golden(name)
    order = int(g["order"])
    binary = refdrv.dropin_binary(order)
    if not os.path.exists(binary):
        pytest.skip("drop-in binary not built (make -C oracle dropin)")
    print(f"[gpu-dropin] {name}", flush=True)
    ns = golden_nodeset(g)
    out, _ = refdrv.run_reference(ns, order, taylor_green_ic(ns), nsteps=int(g["nsteps"]),
                                  phys=golden_phys(g), binary=binary, timeout=300)
    n = ns.npfb
    assert np.array_equal(out["x"][n:], g["gx"]) and np.array_equal(out["irel"], g["irel"])
    assert np.array_equal(out["ij_count"], g["ij_count"])
    x, y = out["x"], out["y"]
    ptr, col, (wx, wl, wh) = csr_from_rows(out["ij_count"], out["ij_link"], x, y, x[:n], y[:n],
                                           [out["w_grad"][:, :, 0], out["w_lap"], out["w_hyp"]])
    assert np.array_equal(col, g["col"])
    assert per_node_rel(ptr, wx, g["wx"]) < WEIGHT_RTOL
    assert per_node_rel(ptr, wl, g["wl"]) < WEIGHT_RTOL
    assert per_node_rel(ptr, wh, g["wh"]) < WEIGHT_RTOL
    for k in ("ro", "rou", "rov", "pxx", "pxy", "pyy", "cxx", "cxy", "cyy"):
        assert field_rel(out[k], g[f"f_{k}"]) < FIELD_RTOL, k
    assert out["itime"] == int(g["itime"])


@pytest.mark.parametrize("name", ["m4", "m6"])
def test_dropin_adapt_stencils_matches_golden(name):
    """adapt_stencils in the reference's own call sequence (find_neighbours,
    adapt_stencils, refine_and_finalise_domain's list release, find_neighbours
    again): h, h_small and the rebuilt lists as the oracle in the device's
    arithmetic (correctly rounded Wendland power) leaves them; the reference's
    own selections differ where libm's pow rounds differently
    (tests/test_gpu_adapt.py bounds that)."""
    import sys
    from parity_util import GOLDEN
    sys.path.insert(0, GOLDEN)
    from make_golden import ACASES, adapt_case
    _, kind, forder, morder, n, hos, seed = {c[0]: c for c in ACASES}[name]
    binary = refdrv.dropin_binary(forder)
    if not os.path.exists(binary):
        pytest.skip("drop-in binary not built (make -C oracle dropin)")
    print(f"[gpu-dropin] adapt {name}", flush=True)
    import oracle_lib as OL
    ns = adapt_case(kind, forder, n, hos, seed)
    out, _ = refdrv.run_reference(ns, forder, taylor_green_ic(ns), nsteps=0, binary=binary,
                                  adapt=True, timeout=300)
    with OL.wab_pow4(1):
        o = OL.OracleSim(ns, morder, forder)
        o.ghosts()
        o.neighbours()
        o.adapt_stencils()
        h, hs = o.h_hs()
        o.neighbours()
    cnt, cs, _ = o.lists()
    assert np.array_equal(out["adapt_h"], h)
    assert np.array_equal(out["adapt_h_small"], hs)
    assert np.array_equal(out["ij_count"], cnt)
    assert np.array_equal(out["ij_count_small"], cs)


def _check_fields(out, g, keys=("ro", "rou", "rov", "pxx", "pxy", "pyy", "cxx", "cxy", "cyy"),
                  conf_rtol=FIELD_RTOL):
    for k in keys:
        err = field_rel(out[k], g[f"f_{k}"])
        print(f"  {k}: {err:.2e}")
        assert err < (FIELD_RTOL if k in ("ro", "rou", "rov") else conf_rtol), k
    for q, name in enumerate(("dt", "time")):
        assert abs(out[name] - g["ctl"][q]) <= 1e-10 * abs(g["ctl"][q]), name
    assert out["itime"] == int(g["itime"])


@pytest.mark.parametrize("name", ["m4_h18", "m6_20"])
@pytest.mark.parametrize("kind", ["dropstep", "droprhs"])
def test_dropin_step_and_rhs_match_golden(kind, name):
    """The reference's own time loop (ref_driver: set_tstep, set_tstep_PID,
    step_rk3_4S_2R_EE) with the step module replaced (the device's fused
    step, state device-resident) or the rhs module replaced (the reference's
    step calling the device's calc_all_rhs and filter_variables)."""
    g = load_golden(name)
    order = int(g["order"])
    binary = os.path.join(refdrv.REF_DIR, f"sunset_{kind}_lc_m{order}")
    if not os.path.exists(binary):
        pytest.skip("drop-in binary not built (make -C oracle dropin)")
    print(f"[gpu-dropin] {kind} {name}", flush=True)
    ns = golden_nodeset(g)
    out, _ = refdrv.run_reference(ns, order, taylor_green_ic(ns), nsteps=int(g["nsteps"]),
                                  phys=golden_phys(g), binary=binary, timeout=300,
                                  fields_only=True)
    _check_fields(out, g)


@pytest.mark.parametrize("name", ["wall_m6", "io_m6_soft", "io_m4_ramp"])
def test_dropin_step_with_boundaries_matches_golden(name):
    """The step drop-in on node sets with walls, or an inflow and an outflow:
    characteristic boundary conditions and the inflow ramp run in the device
    step inside the reference's time loop."""
    from parity_util import golden_bcase
    g, ns, ic, phys, drive, inflow = golden_bcase(name)
    order = int(g["order"])
    binary = os.path.join(refdrv.REF_DIR, f"sunset_dropstep_lc_m{order}")
    if not os.path.exists(binary):
        pytest.skip("drop-in binary not built (make -C oracle dropin)")
    print(f"[gpu-dropin] dropstep {name}", flush=True)
    out, _ = refdrv.run_reference(ns, order, ic, nsteps=int(g["nsteps"]),
                                  phys=dict(phys, drive=(*drive, 0.0)), inflow=inflow,
                                  binary=binary, timeout=300, fields_only=True)
    # the conformation at 1e-8 on these sets, as the device's own run
    # (tests/test_gpu_bound.py)
    _check_fields(out, g, keys=("ro", "rou", "rov", "pxx", "pxy", "pyy"), conf_rtol=1e-8)


@pytest.mark.parametrize("kind", ["wall", "io"])
def test_dropin_laplacian_transverse_matches_reference(kind):
    """calc_laplacian_transverse_only_on_bound and calc_divergence through
    the drop-in derivatives module (the device's tangential boundary weights;
    the wall divergence from boundary-frame normal derivatives) against the
    reference's own routines, on a channel with walls or with an inflow and an
    outflow (ref_driver flag 32: the operator on ro and on cxx, the divergence
    of (rou, rov))."""
    from sunset_amd.nodes import inflow_outflow_channel, wall_channel
    binary = refdrv.dropin_binary(6)
    if not (os.path.exists(binary) and refdrv.available(6)):
        pytest.skip("drop-in or reference binary not built (make -C oracle ref dropin)")
    print(f"[gpu-dropin] transverse Laplacian {kind}", flush=True)
    ns = (wall_channel(16, 16, order=6, seed=1) if kind == "wall"
          else inflow_outflow_channel(16, 16, order=6, seed=2))
    n = ns.npfb
    ic = dict(ro=1.0 + 0.05 * np.sin(2 * np.pi * ns.x) * np.cos(np.pi * ns.y) + 0.02 * ns.y ** 2,
              u=0.1 * np.sin(2 * np.pi * ns.x) * (1.0 + ns.y ** 2),
              v=0.05 * np.cos(2 * np.pi * ns.x) * np.sin(np.pi * ns.y) + 0.02 * ns.y,
              cxx=1.0 + 0.1 * np.cos(2 * np.pi * ns.x) * ns.y, cxy=np.zeros(n), cyy=np.ones(n))
    R, _ = refdrv.run_reference(ns, 6, ic, nsteps=0, transverse=True, timeout=300)
    D, _ = refdrv.run_reference(ns, 6, ic, nsteps=0, transverse=True, binary=binary,
                                timeout=300)
    bnd = (ns.node_type >= 0) & (ns.node_type <= 2)
    for key in ("lt_ro", "lt_cxx", "div_rouv"):
        scale = np.abs(R[key]).max()
        err = np.abs(D[key] - R[key]).max() / scale
        err_b = np.abs(D[key][bnd] - R[key][bnd]).max() / scale
        print(f"  {key}: {err:.2e} (boundary nodes {err_b:.2e})")
        assert err < 1e-9, key

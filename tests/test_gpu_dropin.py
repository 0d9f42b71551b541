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

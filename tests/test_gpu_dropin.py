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

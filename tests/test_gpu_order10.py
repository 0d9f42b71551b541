"""LABFM order 10 on the device (65 unknowns; common_parameter.F90:51-53,
labf.F90:56,66, hovs 3.1): the block-per-system weight solve (k_weights_big)
and the time loop, through the C ABI.

At order 10 the moment systems have cond(A) ~ 5e14: the reference's SVD
weights are within ~1e-4 of the exact solution of its FP64 system, the
oracle's within ~1e-9 (tests/test_oracle_vs_ref.py), the device's LU with a
double-double refinement step within 1e-8 (checked here against 50-digit
solves).  Lists are bit-exact against the reference; weights are compared
with the oracle's per node (max_k |dw| / max_k |w| < 1e-7, both solvers'
error combined); the fields after 10 RK3 steps with the reference's run to
1e-8 (the solver errors lie in directions the smooth fields do not excite)."""
import numpy as np
import pytest

import oracle_lib as OL
from parity_util import (csr_from_rows, field_rel, golden_nodeset, golden_phys, load_golden,
                         per_node_rel, taylor_green_ic)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    import sunset_amd
    print("[gpu-order10] 20^2 lattice, order 10", flush=True)
    g = load_golden("m10_20")
    ns = golden_nodeset(g)
    s = sunset_amd.Solver(morder=10, form="lc").setup(ns)
    cnt, cs, link = s.find_neighbours()
    yield g, ns, s, (cnt, cs, link)
    s.close()


def test_lists_match_reference(case):
    g, ns, s, (cnt, cs, link) = case
    x, y, _, _ = s.ghosts()
    n = ns.npfb
    ptr, col, _ = csr_from_rows(cnt, link, x, y, x[:n], y[:n])
    assert np.array_equal(ptr, g["ptr"]) and np.array_equal(col, g["col"])


def test_weights_vs_extended_precision_and_oracle(case):
    pytest.importorskip("mpmath")
    from parity_util import exact_weights_dx
    g, ns, s, (cnt, cs, link) = case
    x, y, _, _ = s.ghosts()
    n = ns.npfb
    W = s.weights(link.shape[1])
    for i in (0, 137, 399):
        wex = exact_weights_dx(ns, x, y, link, cs, i, 10, dps=50)
        wd = W["w_grad"][i, :cs[i], 0]
        err = np.abs(wd - wex).max() / np.abs(wex).max()
        print(f"node {i}: device vs 50-digit solve {err:.2e}")
        assert err < 1e-8
    o = OL.OracleSim(ns, 10).setup()
    Wo = o.weight_arrays()
    oc, _, ol = o.lists()
    for key, comp in (("w_grad", 0), ("w_grad", 1), ("w_lap", None), ("w_hyp", None)):
        a = W[key][:, :, comp] if comp is not None else W[key]
        b = Wo[key][:, :, comp] if comp is not None else Wo[key]
        ptr, _, (da,) = csr_from_rows(cnt, link, x, y, x[:n], y[:n], [a])
        _, _, (db,) = csr_from_rows(oc, ol, x, y, x[:n], y[:n], [b])
        err = per_node_rel(ptr, da, db)
        print(f"{key}{'' if comp is None else comp}: device vs oracle {err:.2e}")
        assert err < 1e-7
    assert np.allclose(W["vol"], g["vol"], rtol=1e-13, atol=0)


def test_fields_after_steps_match_reference(case):
    g, ns, s, _ = case
    s.set_phys(**golden_phys(g))
    s.set_initial(taylor_green_ic(ns))
    s.run(int(g["nsteps"]))
    F = s.fields()
    assert F["dt"] == pytest.approx(float(g["ctl"][0]), rel=1e-12)
    assert F["itime"] == int(g["itime"])
    for k in ("ro", "rou", "rov", "pxx", "pxy", "pyy", "cxx", "cxy", "cyy"):
        err = field_rel(F[k], g[f"f_{k}"])
        print(f"{k}: {err:.2e}")
        assert err < 1e-8, k


def test_order10_refusals():
    import sunset_amd
    from sunset_amd.nodes import jittered_lattice, wall_channel
    ns = jittered_lattice(12, seed=1, h_over_s=3.1)
    s = sunset_amd.Solver(morder=10, form="lc")
    s.set_nodes(ns)
    s.find_neighbours(copy=False)
    with pytest.raises(sunset_amd.SunsetError, match="order 10"):
        s.adapt_stencils()
    s.close()
    wn = wall_channel(12, 12, order=10, seed=1)
    s = sunset_amd.Solver(morder=10, form="lc")
    s.set_nodes(wn)
    s.find_neighbours(copy=False)
    try:
        s.calc_weights()   # rows near the walls may be singular at this size
    except sunset_amd.SunsetError:
        pass
    with pytest.raises(sunset_amd.SunsetError, match="order 10"):
        s.calc_boundary_weights()
    s.close()

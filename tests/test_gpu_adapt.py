"""Stencil adaptation on the device (sv_adapt_stencils, labf.F90:1079-1721)
through the C ABI.

The adaptation's residual test compares an SVD residual with a threshold at
the SVD's own rounding level, so its decisions follow the last bit of every
input.  The device runs the Golub-Reinsch SVD in the reference's operation
order and evaluates the Wendland kernel's (1 - q/2)^4 correctly rounded; the
reference takes it from libm's pow, which is misrounded on ~0.1% of
arguments.  Parity is therefore stated at two levels:
  * bit for bit against the oracle evaluating the same arithmetic
    (oracle_lib.wab_pow4(1)): h, h_small, reductions and the rebuilt lists,
    on the committed fixtures' node sets and on 96^2 lattices;
  * against the reference (fixtures from the flang build): >= 90% of the
    nodes get the reference's h and h_small, every adapted stencil passes the
    reference's amplification checks, and >= 95% pass its residual check, the
    rest by <= 3x (the reference's own selections fare the same under the
    correctly rounded arithmetic; tests/test_oracle_adapt.py).
The weights on the adapted stencils match the reference's within 1e-10 where
the stencils agree."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as OL
from parity_util import GOLDEN, WEIGHT_RTOL, csr_from_rows, per_node_rel

sys.path.insert(0, GOLDEN)
from make_golden import ACASES, adapt_case  # noqa: E402

pytestmark = pytest.mark.gpu

F98 = float(np.float32(0.98))


@pytest.fixture(scope="module")
def sunset():
    import sunset_amd
    return sunset_amd


def _device_adapt(sunset, ns, morder, forder):
    s = sunset.Solver(morder=morder, forder=forder, form="lc")
    s.set_nodes(ns)
    s.find_neighbours(copy=False)
    h, hs, red = s.adapt_stencils()
    return s, h, hs, red


def _oracle_adapt(ns, morder, forder, mode):
    with OL.wab_pow4(mode):
        o = OL.OracleSim(ns, morder, forder)
        o.ghosts()
        o.neighbours()
        red = o.adapt_stencils()
        h, hs = o.h_hs()
        o.neighbours()
    return o, h, hs, red


@pytest.mark.parametrize("name", [c[0] for c in ACASES])
def test_adapt_matches_oracle_and_reference_fixture(sunset, name):
    print(f"[gpu-adapt] fixture {name}", flush=True)
    _, kind, forder, morder, n, hos, seed = {c[0]: c for c in ACASES}[name]
    g = dict(np.load(os.path.join(GOLDEN, f"ref_adapt_{name}.npz")))
    ns = adapt_case(kind, forder, n, hos, seed)
    s, h, hs, red = _device_adapt(sunset, ns, morder, forder)
    # bit for bit: the oracle in the device's arithmetic
    o, oh, ohs, ored = _oracle_adapt(ns, morder, forder, 1)
    assert np.array_equal(h, oh), f"h differs on {np.count_nonzero(h != oh)} nodes"
    assert np.array_equal(hs, ohs), f"h_small differs on {np.count_nonzero(hs != ohs)} nodes"
    assert np.array_equal(red, ored)
    cnt, cs, link = s.find_neighbours()
    oc, ocs, ol = o.lists()
    assert np.array_equal(cnt, oc) and np.array_equal(cs, ocs)
    for i in range(ns.npfb):
        assert np.array_equal(link[i, :cnt[i]], ol[i, :oc[i]])
    # the reference: agreement, and its own acceptance checks on the device's stencils
    eq = (h == g["h"]) & (hs == g["h_small"])
    print(f"h, h_small equal to the reference's on {eq.mean():.4f} of {ns.npfb} nodes")
    assert eq.mean() >= 0.9
    ref = OL.OracleSim(ns, morder, forder)
    ref.ghosts()
    ref.neighbours()
    for p, hsel in ((1, h), (2, hs)):
        ratio, amp = ref.adapt_check(hsel, p)
        adapted = red[p - 1] > 0
        if adapted.any():
            print(f"pass {p}: {adapted.sum()} adapted, residual over tolerance on "
                  f"{(ratio[adapted] >= 1).sum()} (max {ratio[adapted].max():.2f}x)")
            assert (amp[adapted] == 0).all()
            assert (ratio[adapted] >= 1.0).mean() <= 0.05 and ratio[adapted].max() < 3.0
    if "wx" in g:
        s.calc_weights()
        s.calc_boundary_weights()
        s.calc_sums()
        s.filter_coefficients()
        x, y, _, _ = s.ghosts()
        W = s.weights(link.shape[1])
        nn = ns.npfb
        ptr, col, (wx, wy, wl, wh) = csr_from_rows(
            cnt, link, x, y, x[:nn], y[:nn],
            [W["w_grad"][:, :, 0], W["w_grad"][:, :, 1], W["w_lap"], W["w_hyp"]])
        same = eq & (cnt == g["ij_count"]) & (cs == g["ij_count_small"])
        gp = g["ptr"]
        for a, b, nm in ((wx, g["wx"], "x"), (wy, g["wy"], "y"), (wl, g["wl"], "lap"),
                         (wh, g["wh"], "hyp")):
            worst = 0.0
            for i in np.nonzero(same)[0]:
                assert np.array_equal(col[ptr[i]:ptr[i + 1]], g["col"][gp[i]:gp[i + 1]])
                d = np.abs(a[ptr[i]:ptr[i + 1]] - b[gp[i]:gp[i + 1]]).max()
                worst = max(worst, d / np.abs(b[gp[i]:gp[i + 1]]).max())
            print(f"adapted-stencil weights {nm}: max per-node rel {worst:.3e} "
                  f"on {same.sum()} nodes")
            assert worst < WEIGHT_RTOL, nm


@pytest.mark.parametrize("forder,morder,hos", [(6, 6, 2.4), (8, 6, 2.7)])
def test_adapt_matches_oracle_96(sunset, forder, morder, hos):
    from sunset_amd.nodes import jittered_lattice
    print(f"[gpu-adapt] 96^2 forder {forder} morder {morder}", flush=True)
    ns = jittered_lattice(96, seed=77, h_over_s=hos)
    _, h, hs, red = _device_adapt(sunset, ns, morder, forder)
    _, oh, ohs, ored = _oracle_adapt(ns, morder, forder, 1)
    print(f"mean reductions {red[0].mean():.2f} / {red[1].mean():.2f}")
    assert np.array_equal(h, oh), f"h differs on {np.count_nonzero(h != oh)} nodes"
    assert np.array_equal(hs, ohs), f"h_small differs on {np.count_nonzero(hs != ohs)} nodes"
    assert np.array_equal(red, ored)


def test_adapt_refusals(sunset):
    from sunset_amd.nodes import jittered_lattice, taylor_green_ic
    ns = jittered_lattice(16, seed=5, h_over_s=2.4)
    s = sunset.Solver(morder=6, forder=4, form="lc")
    s.set_nodes(ns)
    s.find_neighbours(copy=False)
    with pytest.raises(sunset.SunsetError, match="morder <= forder"):
        s.adapt_stencils()
    s = sunset.Solver(morder=6, form="lc")
    with pytest.raises(sunset.SunsetError, match="sv_find_neighbours"):
        s.set_nodes(ns).adapt_stencils()
    s.setup(ns)
    s.set_phys()
    s.set_initial(taylor_green_ic(ns))
    with pytest.raises(sunset.SunsetError, match="before sv_set_initial"):
        s.adapt_stencils()

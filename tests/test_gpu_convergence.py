"""Convergence order of the derivative operators on an analytic field equals
the reference's: rho = 1 + 0.1 sin(2 pi x) cos(2 pi y) on seeded periodic
jittered lattices of 16^2, 32^2 and 64^2 nodes, orders 4 (h = 1.8 s) and 6
(h = 2.4 s).  The reference's maximum errors and observed orders are the
committed vectors tests/golden/ref_conv_m{4,6}.npz (written by
tests/golden/make_golden.py from the compiled reference's calc_gradient and
calc_laplacian).  The device's errors must equal the reference's to 1e-6
relative or 1e-9 absolute (same lists, weights equal to 1e-10) and its orders
to 0.01."""
import os

import numpy as np
import pytest

from parity_util import GOLDEN
from sunset_amd.nodes import jittered_lattice

pytestmark = pytest.mark.gpu


def _field(x, y):
    return 1.0 + 0.1 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)


def _exact(x, y):
    k = 2 * np.pi
    return (0.1 * k * np.cos(k * x) * np.cos(k * y), -0.1 * k * np.sin(k * x) * np.sin(k * y),
            -2 * k * k * 0.1 * np.sin(k * x) * np.cos(k * y))


@pytest.mark.parametrize("order", [4, 6])
def test_convergence_order_equals_reference(order):
    import sunset_amd
    g = np.load(os.path.join(GOLDEN, f"ref_conv_m{order}.npz"))
    errs = []
    for n in g["sizes"]:
        ns = jittered_lattice(int(n), seed=int(g["seed"]), h_over_s=float(g["h_over_s"]))
        print(f"[conv] order {order} n {n}", flush=True)
        s = sunset_amd.Solver(morder=order, form="lc").setup(ns)
        x, y, _, _ = s.ghosts()
        f = _field(x, y)     # ghosts are periodic images: the field is periodic
        gx, gy = s.gradient(f)
        lap = s.laplacian(f)
        s.close()
        ex, ey, el = _exact(ns.x, ns.y)
        errs.append([np.abs(gx - ex).max(), np.abs(gy - ey).max(), np.abs(lap - el).max()])
    errs = np.array(errs)
    rates = np.log2(errs[:-1] / errs[1:])
    # weights equal to ~1e-11 relative show in the smallest errors as ~1e-10
    np.testing.assert_allclose(errs, g["errs"], rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(rates, g["rates"], atol=0.01)
    # the orders the method is built for: m for gradients, ~m - 1 for Laplacians
    assert rates[-1, :2].min() > order - 0.5 and rates[-1, 2] > order - 1.6

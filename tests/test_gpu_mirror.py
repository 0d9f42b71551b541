"""Symmetric and no-slip mirror edges (mirror_boundaries.F90:34-352): ghosts
are mirror images whose velocity and conformation xy component change sign
(vrelation 2: x mirror, 3: y mirror, 4: no-slip or a corner of two mirrors).
The device folds every ghost link to its parent and applies the sign flips
when it stages the block's neighbour records; RK3 steps against the oracle,
whose mirror handling is checked against the compiled reference in
test_oracle_vs_ref.test_mirror_edges_vs_reference."""
import numpy as np
import pytest

import oracle_lib as OL
from parity_util import FIELD_RTOL, field_rel
from sunset_amd.nodes import jittered_lattice

pytestmark = pytest.mark.gpu

PH = dict(Re=10.0, Wi=0.5, beta=0.5, eps=0.1, Ma=0.1)


def mirror_ic(ns):
    n = ns.npfb
    x, y = ns.x, ns.y
    return dict(ro=1 + 0.01 * np.cos(2 * np.pi * x), u=0.1 * np.sin(np.pi * x) * np.cos(2 * np.pi * y),
                v=0.1 * np.sin(np.pi * y) * np.cos(np.pi * x),
                cxx=1 + 0.05 * np.cos(np.pi * x) ** 2, cxy=0.02 * np.sin(np.pi * x) * np.sin(np.pi * y),
                cyy=np.ones(n))


@pytest.mark.parametrize("form", ["lc", "di"])
@pytest.mark.parametrize("bc", [(2, 2, 1, 1), (1, 1, 2, 2), (2, 2, 2, 2), (1, 1, 3, 3)])
def test_mirror_edges_vs_oracle(bc, form):
    import sunset_amd
    ns = jittered_lattice(20, seed=5, h_over_s=2.4, bc=bc)
    ic = mirror_ic(ns)
    print(f"[mirror] bc {bc} form {form}", flush=True)
    s = sunset_amd.Solver(morder=6, form=form).setup(ns)
    s.set_phys(**PH)
    s.set_initial(ic)
    s.run(20)
    F = s.fields()
    s.close()
    # no-slip edges give nodes near the wall up to ~360 neighbours (periodic
    # images of the mirror images), beyond the reference's default nplink 200
    o = OL.OracleSim(ns, 6, nplink=600).setup()
    o.make_phys(form=form, **PH)
    o.set_initial(ic)
    o.run(20)
    R = o.fields()
    assert F["dt"] == pytest.approx(R["dt"], rel=1e-12)
    keys = ("ro", "rou", "rov") + (("pxx", "pxy", "pyy") if form == "lc" else ("cxx", "cxy", "cyy"))
    for k in keys:
        assert field_rel(F[k], R[k]) < FIELD_RTOL, k


@pytest.mark.parametrize("bc", [(2, 2, 1, 1), (1, 1, 3, 3)])
def test_mirror_direct_form_polymer_diffusion_vs_oracle(bc):
    """Polymer diffusion (the Laplacian of c) across mirror edges: the mirrored
    records carry the negated xy component into the Laplacian as into the
    gradients."""
    import sunset_amd
    # (the seed-6 lattice of this size with no-slip edges diverges within 7
    # steps in the direct form, in the oracle as on the device, with or
    # without diffusion; the seed-5 one is stable)
    ns = jittered_lattice(20, seed=5, h_over_s=2.4, bc=bc)
    ic = mirror_ic(ns)
    ph = dict(PH, Mdiff=2e-3)
    print(f"[mirror] bc {bc} DI Mdiff", flush=True)
    s = sunset_amd.Solver(morder=6, form="di").setup(ns)
    s.set_phys(**ph)
    s.set_initial(ic)
    s.run(20)
    F = s.fields()
    s.close()
    o = OL.OracleSim(ns, 6, nplink=600).setup()
    o.make_phys(form="di", **ph)
    o.set_initial(ic)
    o.run(20)
    R = o.fields()
    assert F["dt"] == pytest.approx(R["dt"], rel=1e-12)
    for k in ("ro", "rou", "rov", "cxx", "cxy", "cyy"):
        assert field_rel(F[k], R[k]) < FIELD_RTOL, k

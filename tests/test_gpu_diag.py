"""On-device diagnostics (sv_diagnostics, statistics.F90:305-654) against the
oracle's restatement (pinned by closed forms in test_oracle_diag.py) on the
same states: after LC and DI runs, on the analytic Kolmogorov and
Taylor-Green fields, and per ensemble member (bitwise equal to the member's
own context)."""
import numpy as np
import pytest

import oracle_lib as OL
from sunset_amd.nodes import jittered_lattice, taylor_green_ic

pytestmark = pytest.mark.gpu

PH = dict(Re=1.0, Wi=2.0, beta=0.1, eps=0.25, Ma=0.1, kol_amp=1.0, kol_k=4 * np.pi)


def _pair(ns, form, ph, ic, order=6, dt=None):
    import sunset_amd
    s = sunset_amd.Solver(morder=order, form=form).setup(ns)
    s.set_phys(**ph)
    if dt is None:
        s.set_initial(ic)
    else:
        s.set_initial(ic, dt=dt, dt_mode="fixed")
    o = OL.OracleSim(ns, order).setup()
    o.make_phys(form=form, dt_mode="pid" if dt is None else "fixed", **ph)
    o.set_initial(ic, dt=dt)
    return s, o


@pytest.mark.parametrize("form", ["lc", "di"])
def test_mass_energy_and_conf_after_steps(form):
    ns = jittered_lattice(32, seed=5, h_over_s=2.4)
    print(f"[diag] {form} 32^2 order 6, 6 steps", flush=True)
    s, o = _pair(ns, form, PH, taylor_green_ic(ns))
    s.run(6)
    o.run(6)
    for kind in ("mass_energy", "conf"):
        d, r = s.diagnostics(kind), o.diagnostics(kind)
        np.testing.assert_allclose(d, r, rtol=1e-9, atol=1e-13, err_msg=kind)
    s.close()


def test_kolmogorov_and_taylor_green_fields():
    ns = jittered_lattice(32, seed=6, h_over_s=2.4)
    Wi = 0.7
    y = ns.y
    s2 = np.sin(2 * y)
    ic = dict(ro=np.ones(ns.npfb), u=np.cos(2 * y) + 1e-3, v=np.zeros(ns.npfb),
              cxx=1 + 8 * Wi * Wi * s2 * s2, cxy=-2 * Wi * s2, cyy=np.ones(ns.npfb))
    print("[diag] kolmogorov profile", flush=True)
    s, o = _pair(ns, "di", dict(PH, Wi=Wi, eps=0.0), ic, dt=1e-4)
    np.testing.assert_allclose(s.diagnostics("kolmogorov"), o.diagnostics("kolmogorov"),
                               rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(s.diagnostics("kolmogorov"), [1e-3, 0, 0, 0, 0], rtol=1e-12,
                               atol=1e-15)
    s.close()
    print("[diag] taylor-green decay, 10 steps", flush=True)
    x = ns.x
    ic = dict(ro=np.ones(ns.npfb), u=-np.cos(2 * np.pi * x) * np.sin(2 * np.pi * y),
              v=np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y), cxx=np.ones(ns.npfb),
              cxy=np.zeros(ns.npfb), cyy=np.ones(ns.npfb))
    s, o = _pair(ns, "di", dict(Re=10.0, Wi=1.0, beta=0.9, eps=0.0, Ma=0.1), ic, dt=2e-5)
    assert s.diagnostics("tg")[0] < 1e-14
    s.run(10)
    o.run(10)
    d, r = s.diagnostics("tg")[0], o.diagnostics("tg")[0]
    assert abs(d - r) <= 1e-8 * abs(r) + 1e-15, (d, r)
    s.close()


def test_ensemble_member_diagnostics_equal_solo_contexts():
    import sunset_amd
    sizes = [(20, 4), (28, 6)]
    solos, members = [], []
    for k, (n, order) in enumerate(sizes):
        ns = jittered_lattice(n, seed=10 + k, h_over_s=2.4 if order == 6 else 1.8)
        for lst in (solos, members):
            s = sunset_amd.Solver(morder=order, form="lc").setup(ns)
            s.set_phys(**dict(PH, Wi=1.0 + k))
            s.set_initial(taylor_green_ic(ns))
            lst.append(s)
    print("[diag] ensemble of 2, 5 steps", flush=True)
    E = sunset_amd.Ensemble(members)
    E.run(5)
    for m, s in enumerate(solos):
        s.run(5)
        for kind in ("mass_energy", "conf"):
            assert np.array_equal(E.diagnostics(m, kind), s.diagnostics(kind)), (m, kind)
    with pytest.raises(sunset_amd.SunsetError):
        E.diagnostics(0, "kolmogorov")
    E.close()
    for s in solos + members:
        s.close()

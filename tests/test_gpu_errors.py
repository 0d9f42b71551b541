"""Error paths and small/ragged inputs through the C ABI on the device.

The reference stops (write + stop / MPI_Abort) where these entry points
return a negative SV_E_* code with a message; every case here is turned into
an error on the host side before any kernel runs.  The ragged cases run the
whole path (neighbours, weights, RK steps) on node counts that fill no full
256-node block or 64-node slice, against the oracle."""
import numpy as np
import pytest

import oracle_lib as OL
from sunset_amd.nodes import jittered_lattice, taylor_green_ic

pytestmark = pytest.mark.gpu

PH = dict(Re=1.0, Wi=1.0, beta=0.5, eps=0.0, Ma=0.1)


def _sv():
    import sunset_amd
    return sunset_amd


def _code(exc):
    return int(str(exc.value).split(":")[0].split()[-1])


def test_call_order_and_argument_errors():
    sv = _sv()
    ns = jittered_lattice(12, seed=1, h_over_s=1.8)
    s = sv.Solver(morder=4, form="lc")
    print("[errors] call order", flush=True)
    with pytest.raises(sv.SunsetError) as e:          # null node arrays
        s._ck(sv.lib().sv_set_nodes(s.h, 4, None, None, None, None, None, None))
    assert _code(e) == -1
    bad = jittered_lattice(12, seed=1, h_over_s=1.8)
    bad.h_small = bad.h * 1.1                        # h_small > h
    with pytest.raises(sv.SunsetError, match="h_small"):
        s.set_nodes(bad)
    s.set_nodes(ns)
    with pytest.raises(sv.SunsetError):              # weights before neighbours
        s.calc_weights()
    with pytest.raises(sv.SunsetError) as e:          # list longer than nplink
        s.find_neighbours(nplink=4)
    assert _code(e) == -4
    s.find_neighbours()
    with pytest.raises(sv.SunsetError):
        s.calc_weights(variant=7)                    # no such launch variant
    s.calc_weights()
    s.calc_sums()
    s.filter_coefficients()
    s.set_phys(**PH)
    with pytest.raises(sv.SunsetError):              # steps before an initial state
        s.run(1)
    ic = taylor_green_ic(ns)
    with pytest.raises(sv.SunsetError, match="density"):
        s.set_initial(dict(ic, ro=-ic["ro"]))
    with pytest.raises(sv.SunsetError, match="positive definite"):
        s.set_initial(dict(ic, cxy=np.full(ns.npfb, 2.0)))
    s.set_initial(ic)
    with pytest.raises(sv.SunsetError):              # the adjoint is DI-only
        s.adj_setup(5)
    with pytest.raises(sv.SunsetError):
        s._ck(sv.lib().sv_diagnostics(s.h, 9, 0, None))
    with pytest.raises(sv.SunsetError):
        s._ck(sv.lib().sv_diagnostics(s.h, 0, 3, np.zeros(3).ctypes.data_as(sv.core.PD)))
    s.run(2)
    assert np.isfinite(s.fields()["ro"]).all()       # the context is still usable
    s.close()


def test_adjoint_call_order_errors():
    sv = _sv()
    ns = jittered_lattice(12, seed=2, h_over_s=1.8)
    s = sv.Solver(morder=4, form="di").setup(ns)
    s.set_phys(**PH)
    s.set_initial(taylor_green_ic(ns))               # PID time step
    with pytest.raises(sv.SunsetError, match="fixed time step"):
        s.adj_setup(5)
    s.set_initial(taylor_green_ic(ns), dt=1e-4, dt_mode="fixed")
    s.adj_setup(5)
    with pytest.raises(sv.SunsetError, match="sv_rollout_forward"):
        s.rollout_vjp(1.0)
    with pytest.raises(sv.SunsetError):
        s.adj_setup(0)                               # checkpoint interval out of range
    s.close()


def test_ensemble_rejects_mixed_forms():
    sv = _sv()
    members = []
    for k, form in enumerate(("lc", "di")):
        ns = jittered_lattice(10, seed=3 + k, h_over_s=1.8)
        s = sv.Solver(morder=4, form=form).setup(ns)
        s.set_phys(**PH)
        s.set_initial(taylor_green_ic(ns))
        members.append(s)
    with pytest.raises(sv.SunsetError, match="member 1: conformation form"):
        sv.Ensemble(members)
    # the message is also left on the first member
    assert "conformation form" in sv.lib().sv_last_error(members[0].h).decode()
    for s in members:
        s.close()


def test_ensemble_refuses_mirror_member_beside_member_without_halos():
    """Mirror flips live in the per-block halo keys, so an ensemble with a
    mirror-ghost member needs halos in every member; a member built with
    halo_sweeps off is refused with a message."""
    sv = _sv()
    members = []
    for k, (bc, halo) in enumerate((((2, 2, 1, 1), 1), ((1, 1, 1, 1), 0))):
        ns = jittered_lattice(12, seed=7 + k, h_over_s=1.8, bc=bc)
        s = sv.Solver(morder=4, form="lc").setup(ns)
        s.set_option("halo_sweeps", halo)
        s.set_phys(**PH)
        s.set_initial(taylor_green_ic(ns))
        members.append(s)
    with pytest.raises(sv.SunsetError, match="mirror ghosts"):
        sv.Ensemble(members)
    for s in members:
        s.close()


@pytest.mark.parametrize("nx,ny,order", [(7, 9, 4), (17, 8, 4), (11, 13, 6)])
def test_ragged_node_sets_against_the_oracle(nx, ny, order):
    """63, 136 and 143 nodes: partial 64-node slices and a single partial
    256-node block through neighbours, weights and 3 RK steps."""
    sv = _sv()
    hos = 1.8 if order == 4 else 2.4
    ns = jittered_lattice(nx, ny, seed=9, h_over_s=hos)
    print(f"[errors] ragged {ns.npfb} nodes order {order}", flush=True)
    s = sv.Solver(morder=order, form="lc").setup(ns)
    s.set_phys(**PH)
    s.set_initial(taylor_green_ic(ns))
    s.run(3)
    F = s.fields()
    s.close()
    o = OL.OracleSim(ns, order).setup()
    o.make_phys(form="lc", **PH)
    o.set_initial(taylor_green_ic(ns))
    o.run(3)
    R = o.fields()
    for k in ("ro", "rou", "rov", "pxx", "pxy", "pyy"):
        err = np.abs(F[k] - R[k]).max() / max(np.abs(R[k]).max(), 1.0)
        assert err < 1e-9, (k, err)

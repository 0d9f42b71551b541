"""Ensembles whose members have boundary nodes (walls, soft/hard inflows with a
ramped inflow velocity, outflows) or symmetric / no-slip mirror ghosts, mixed
with periodic members.  Each member's boundary records are concatenated with
its FD-row slots rebased, its mirror flip codes travel in the rebased halo
keys, and its inflow ramp lives in its own control block; every member must end
bit for bit where the same member run alone ends, eager and graph-replayed,
with one launch over all blocks and with one launch per order group."""
import math

import numpy as np
import pytest

from sunset_amd.nodes import inflow_outflow_channel, jittered_lattice, wall_channel

pytestmark = pytest.mark.gpu

PH = dict(Re=10.0, Wi=0.5, beta=0.5, eps=0.1, Ma=0.1)
KEYS = ("ro", "rou", "rov", "u", "v", "pxx", "pxy", "pyy", "cxx", "cxy", "cyy")


def _ic(ns, kind):
    n = ns.npfb
    x, y = ns.x, ns.y
    if kind == "wall":
        ic = dict(ro=np.ones(n), u=4 * y * (1 - y),
                  v=0.01 * np.sin(2 * np.pi * x) * np.sin(np.pi * y) ** 2)
        t = np.minimum(y, 1.0 - y)
    elif kind == "io":
        ic = dict(ro=np.ones(n), u=1.0 + 0.01 * np.sin(2 * np.pi * y),
                  v=0.01 * np.cos(2 * np.pi * y) * np.sin(np.pi * x))
        t = x
    else:
        ic = dict(ro=1 + 0.01 * np.cos(2 * np.pi * x),
                  u=0.1 * np.sin(np.pi * x) * np.cos(2 * np.pi * y),
                  v=0.1 * np.sin(np.pi * y) * np.cos(np.pi * x))
        t = 0.5 * (1 + np.cos(np.pi * x) * np.cos(np.pi * y))
    ic.update(cxx=1.0 + 0.2 * t, cxy=0.05 * t, cyy=1.0 + 0.1 * t ** 2)
    return ic


# (kind, side, order, inflow settings, body force x, Mdiff in the direct form)
MEMBERS = [
    ("wall", 20, 6, {}, 8.0, 2e-3),
    ("io", 18, 4, dict(inflow_type=1, u_control=1, u_start=0.5, u_end=1.0, ramp_time=1e-4),
     0.0, 0.0),
    ("mirror", 20, 6, {}, 0.0, 1e-3),
    ("periodic", 24, 4, {}, 0.0, 0.0),
    ("io", 16, 6, dict(inflow_type=0), 0.0, 2e-3),
]


def _member(sv, kind, side, order, inflow, fx, mdiff, form, seed):
    if kind == "wall":
        ns = wall_channel(side, side, order=order, seed=seed)
    elif kind == "io":
        ns = inflow_outflow_channel(side, side, order=order, seed=seed)
    elif kind == "mirror":
        ns = jittered_lattice(side, seed=seed, h_over_s=2.4, bc=(1, 1, 3, 3))
    else:
        ns = jittered_lattice(side, seed=seed, h_over_s={4: 1.8, 6: 2.4}[order])
    s = sv.Solver(morder=order, form=form).setup(ns)
    if inflow:
        s.set_inflow(**inflow)
    s.set_phys(drive=(fx, 0.0, 0.0), Mdiff=mdiff if form == "di" else 0.0, **PH)
    s.set_initial(_ic(ns, kind))
    return s


@pytest.mark.parametrize("form", ["lc", "di"])
def test_ensemble_with_boundaries_matches_members_alone(form):
    import sunset_amd
    print(f"[ensemble-bound] form {form}, {len(MEMBERS)} members", flush=True)
    steps = 10
    alone = []
    for k, m in enumerate(MEMBERS):
        s = _member(sunset_amd, *m, form=form, seed=k + 1)
        s.run(steps)
        alone.append(s.fields())
        s.close()
    for graph, groups in ((False, 0), (True, 0), (False, 1), (True, 1)):
        print(f"[ensemble-bound] graph {graph} groups {groups}", flush=True)
        members = [_member(sunset_amd, *m, form=form, seed=k + 1) for k, m in enumerate(MEMBERS)]
        E = sunset_amd.Ensemble(members)
        E.set_option("ensemble_groups", groups)
        E.run(steps // 2, graph=graph)
        _ = E.fields(0)
        E.run(steps - steps // 2, graph=graph)
        for k in range(len(MEMBERS)):
            F = E.fields(k)
            assert F["itime"] == steps
            for key in KEYS:
                if form == "di" and key.startswith("p"):
                    continue
                assert np.array_equal(F[key], alone[k][key]), (graph, groups, k, key)
            assert F["dt"] == alone[k]["dt"] and F["time"] == alone[k]["time"], (graph, k)
        E.close()
        for s in members:
            s.close()
    # walls hold zero velocity in the ensemble as alone
    walls0 = alone[0]["rou"]
    assert np.isfinite(walls0).all()
    assert math.isfinite(alone[1]["dt"])

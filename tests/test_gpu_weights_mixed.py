"""Mixed-order weight setup on a cylinder-in-channel node set (order 8 in the
interior, 4 within three local spacings of any wall), through the C ABI.

Small instance (1.5k nodes): neighbour lists bit-exact vs the oracle, weights
within 1e-10 per node (max_k |dw| / max_k |w|) of the oracle's, for the
one-block-per-system launch and the persistent work-queue launch, which must
agree bit for bit.  Full instance (s = 0.004 .. 0.016, ~126k nodes):
size-independent checks -- no singular system, persistent == plain bitwise,
every sampled node reproduces all monomials up to its own order."""
import numpy as np
import pytest

import oracle_lib as OL
from parity_util import WEIGHT_RTOL, per_node_rel

pytestmark = pytest.mark.gpu


def _idx(order):
    return [(d - b, b) for d in range(1, order + 1) for b in range(d + 1)]


def _device_weights(ns, om, persistent):
    import sunset_amd
    s = sunset_amd.Solver(morder=8, form="lc")
    s.set_nodes(ns)
    cnt, cs, link = s.find_neighbours()
    s.calc_weights(order_map=om, persistent=persistent)   # raises on a zero pivot
    s.calc_sums()
    s.filter_coefficients()
    W = s.weights(link.shape[1])
    x, y, _, _ = s.ghosts()
    s.close()
    return cnt, cs, link, W, x, y


def test_small_channel_lists_and_weights_vs_oracle():
    from sunset_amd.nodes import cylinder_channel
    ns, om = cylinder_channel(smin=0.04, smax=0.12, grow=1.0)
    print(f"[mixed] small channel {ns.npfb} nodes", flush=True)
    o = OL.OracleSim(ns, 8)
    o.ghosts()
    o.neighbours()
    o.weights_ordermap(om)
    oc, ocs, ol = o.lists()
    Wo = o.weight_arrays()
    res = {}
    for persistent in (False, True):
        cnt, cs, link, W, x, y = _device_weights(ns, om, persistent)
        assert np.array_equal(cnt, oc) and np.array_equal(cs, ocs)
        for i in range(ns.npfb):
            assert np.array_equal(link[i, : cnt[i]], ol[i, : cnt[i]])
        ptr = np.concatenate([[0], np.cumsum(cnt)])
        pack = lambda w: np.concatenate([w[i, : cnt[i]] for i in range(ns.npfb)])
        for name, a, b in (("x", W["w_grad"][:, :, 0], Wo["w_grad"][:, :, 0]),
                           ("y", W["w_grad"][:, :, 1], Wo["w_grad"][:, :, 1]),
                           ("lap", W["w_lap"], Wo["w_lap"]), ("hyp", W["w_hyp"], Wo["w_hyp"])):
            err = per_node_rel(ptr, pack(a), pack(b[:, : a.shape[1]]))
            print(f"persistent={persistent} weights {name}: max per-node rel {err:.3e}")
            assert err < WEIGHT_RTOL, (persistent, name)
        assert np.allclose(W["vol"], Wo["vol"], rtol=1e-12)
        res[persistent] = W
    for k in ("w_grad", "w_lap", "w_hyp", "vol"):
        assert np.array_equal(res[False][k], res[True][k]), k


@pytest.mark.slow
def test_full_channel_moment_conditions_and_launch_equivalence():
    from sunset_amd.nodes import cylinder_channel
    ns, om = cylinder_channel()
    print(f"[mixed] full channel {ns.npfb} nodes", flush=True)
    out = {}
    for persistent in (False, True):
        out[persistent] = _device_weights(ns, om, persistent)
    cnt, cs, link, W, x, y = out[True]
    for k in ("w_grad", "w_lap", "w_hyp"):
        assert np.array_equal(out[False][3][k], W[k]), k
    rng = np.random.default_rng(5)
    worst = {4: 0.0, 8: 0.0}
    for i in rng.choice(ns.npfb, 400, replace=False):
        m = int(om[i])
        js = link[i, : cnt[i]]
        X, Y = x[js] - x[i], y[js] - y[i]
        for (a, b) in _idx(m):
            mono = X ** a * Y ** b
            for w, exact in ((W["w_grad"][i, : cnt[i], 1], 1.0 if (a, b) == (0, 1) else 0.0),
                             (W["w_lap"][i, : cnt[i]],
                              2.0 if (a, b) in ((2, 0), (0, 2)) else 0.0)):
                scale = np.abs(w).sum() * np.abs(mono).max()
                worst[m] = max(worst[m], abs(w @ mono - exact) / max(scale, 1.0))
    print(f"moment residuals: {worst}")
    assert max(worst.values()) < 1e-9

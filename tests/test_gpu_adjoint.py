"""Device reverse mode (DI, fixed dt) through the C ABI, checked against the
oracle's tangent-linear model (pinned by finite differences in
test_oracle_tlm.py) and against central finite differences of the device's
own forward rollout:
  - right-hand-side VJP:   <rbar, J dU>_oracle == <J^T rbar, dU>_device
  - final-state VJP:       <ubar, dU_T>_oracle == <grad0, dU>_device
  - loss gradient:         dL_oracle == <grad0, dU>_device, and central FD
  - the state after the VJP is the state after the forward rollout, bitwise
  - torch.autograd.Function around the rollout: CPU tensors at the boundary,
    and CUDA tensors with no host copy (a process importing torch first)
Tolerances: 1e-9 relative against the tangent-linear model (rounding of two
different evaluation orders), 1e-6 against finite differences (step 1e-6)."""
import numpy as np
import pytest

import oracle_lib as OL
from sunset_amd.nodes import jittered_lattice, perturbed_rest_ic

pytestmark = pytest.mark.gpu

KEYS = ("ro", "rou", "rov", "cxx", "cxy", "cyy")
PH = dict(Re=1.0, Wi=1.0, beta=0.5, eps=0.1, Ma=0.1, kol_amp=1.0, kol_k=4 * np.pi)
DT = 2e-4


def _ic(ns):
    ic = perturbed_rest_ic(ns, amp=0.05, seed=3)
    ic["cxx"] = 1.0 + 0.2 * np.sin(2 * np.pi * ns.x) ** 2
    ic["cxy"] = 0.1 * np.cos(2 * np.pi * ns.y)
    return ic


def _state(ic):
    return dict(ro=ic["ro"], rou=ic["ro"] * ic["u"], rov=ic["ro"] * ic["v"], cxx=ic["cxx"],
                cxy=ic["cxy"], cyy=ic["cyy"])


def _dir(n, seed):
    rng = np.random.default_rng(seed)
    return {k: rng.standard_normal(n) * (0.01 if k == "ro" else 0.1) for k in KEYS}


def _dot(a, b):
    return sum(float(np.dot(a[k], b[k])) for k in KEYS)


def _device(ns, ph, order=4):
    import sunset_amd
    s = sunset_amd.Solver(morder=order, form="di").setup(ns)
    s.set_phys(**ph)
    return s


def _oracle(ns, ph, ic, order=4):
    o = OL.OracleSim(ns, order).setup()
    o.make_phys(form="di", dt_mode="fixed", **ph)
    o.set_initial(ic, dt=DT)
    return o


@pytest.mark.parametrize("mdiff", [0.0, 2e-3])
def test_rhs_vjp_dot_product_with_oracle_tangent(mdiff):
    ns = jittered_lattice(16, seed=2, h_over_s=1.8)
    ic = _ic(ns)
    ph = dict(PH, Mdiff=mdiff)
    print(f"[adjoint] rhs vjp mdiff={mdiff}", flush=True)
    s = _device(ns, ph)
    s.set_state(_state(ic), dt=DT)
    s.adj_setup(5)
    dU = _dir(ns.npfb, 7)
    rbar = _dir(ns.npfb, 9)
    ub = s.rhs_vjp(rbar)
    tl = _oracle(ns, ph, ic).rhs_tlm(dU)
    tl = dict(zip(KEYS, (tl["ro"], tl["rou"], tl["rov"], tl["xx"], tl["xy"], tl["yy"])))
    lhs, rhs = _dot(rbar, tl), _dot(ub, dU)
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs), (lhs, rhs)
    s.close()


@pytest.mark.parametrize("lanes", [1, 4])
def test_rollout_vjp_matches_tangent_linear_model_and_fd(lanes):
    """Both layouts of the transposed gathers: one lane per node, and four
    lanes per node (links split over a quad, partial sums added in a fixed
    order)."""
    ns = jittered_lattice(16, seed=4, h_over_s=1.8)
    ic = _ic(ns)
    U0 = _state(ic)
    steps, every = 25, 7          # several checkpoint intervals, a short last one
    dU = _dir(ns.npfb, 11)
    print(f"[adjoint] rollout, {lanes} lanes per node", flush=True)
    s = _device(ns, PH)
    s.set_option("adj_lanes", lanes)
    s.set_state(U0, dt=DT)
    s.adj_setup(every)
    L = s.rollout_forward(steps)
    end = s.fields()
    g = s.rollout_vjp(1.0)
    after = s.fields()
    for k in ("ro", "rou", "rov", "cxx", "cxy", "cyy", "u", "v"):
        assert np.array_equal(end[k], after[k]), k
    Lo, dLo, _ = _oracle(ns, PH, ic).rollout_tlm(dU, steps)
    assert abs(L - Lo) <= 1e-10 * abs(Lo) + 1e-15
    dL = _dot(g, dU)
    assert abs(dL - dLo) <= 1e-9 * abs(dLo), (dL, dLo)
    # central differences of the device forward rollout
    eps = 1e-6
    Lpm = []
    for sign in (1, -1):
        s.set_state({k: U0[k] + sign * eps * dU[k] for k in KEYS}, dt=DT)
        Lpm.append(s.rollout_forward(steps))
    fd = (Lpm[0] - Lpm[1]) / (2 * eps)
    assert abs(fd - dL) <= 1e-6 * abs(dL), (fd, dL)
    s.close()


def test_final_state_cotangent_matches_tangent():
    ns = jittered_lattice(16, seed=5, h_over_s=1.8)
    ic = _ic(ns)
    steps = 9
    dU = _dir(ns.npfb, 13)
    ubar = _dir(ns.npfb, 17)
    s = _device(ns, dict(PH, Mdiff=1e-3))
    s.set_state(_state(ic), dt=DT)
    s.adj_setup(4)
    s.rollout_forward(steps)
    g = s.rollout_vjp(0.0, ubar_final=ubar)
    _, _, dUT = _oracle(ns, dict(PH, Mdiff=1e-3), ic).rollout_tlm(dU, steps)
    lhs, rhs = _dot(ubar, dUT), _dot(g, dU)
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs), (lhs, rhs)
    s.close()


def test_torch_autograd_rollout():
    torch = pytest.importorskip("torch")
    from sunset_amd.autograd import rollout_loss
    ns = jittered_lattice(16, seed=6, h_over_s=1.8)
    ic = _ic(ns)
    U0 = np.stack([_state(ic)[k] for k in KEYS])
    d = np.stack([_dir(ns.npfb, 19)[k] for k in KEYS])
    s = _device(ns, PH)
    x = torch.tensor(U0, dtype=torch.float64, requires_grad=True)
    L = rollout_loss(s, x, 12, DT, checkpoint_every=5)
    L.backward()
    dL = float((x.grad * torch.from_numpy(d)).sum())
    eps = 1e-6
    with torch.no_grad():
        Lp = float(rollout_loss(s, torch.from_numpy(U0 + eps * d), 12, DT, 5))
        Lm = float(rollout_loss(s, torch.from_numpy(U0 - eps * d), 12, DT, 5))
    fd = (Lp - Lm) / (2 * eps)
    assert abs(fd - dL) <= 1e-6 * abs(dL), (fd, dL)
    s.close()


def test_rhs_parameter_gradient_matches_oracle_tangent():
    """<rbar, dR/dtheta> for theta = Wi, beta, eps against the oracle's
    parameter tangent (pinned by finite differences in test_oracle_tlm.py)."""
    ns = jittered_lattice(16, seed=2, h_over_s=1.8)
    ic = _ic(ns)
    ph = dict(PH, eps=0.25, Wi=1.3, beta=0.4, Mdiff=1e-3)
    print("[adjoint] rhs parameter gradient", flush=True)
    s = _device(ns, ph)
    s.set_state(_state(ic), dt=DT)
    s.adj_setup(5)
    rbar = _dir(ns.npfb, 9)
    s.rhs_vjp(rbar)
    pg = s.adj_param_grad()
    zero = {k: np.zeros(ns.npfb) for k in KEYS}
    for name in ("Wi", "beta", "eps"):
        tl = _oracle(ns, ph, ic).rhs_tlm(zero, {name: 1.0})
        tl = dict(zip(KEYS, (tl["ro"], tl["rou"], tl["rov"], tl["xx"], tl["xy"], tl["yy"])))
        ref = _dot(rbar, tl)
        assert abs(pg[name] - ref) <= 1e-9 * abs(ref), (name, pg[name], ref)
    s.close()


def test_rollout_parameter_gradient_matches_tangent_and_fd():
    """dL/dWi, dL/dbeta, dL/deps of a 25-step rollout (checkpoints every 7):
    oracle parameter tangent at 1e-9, central differences of the device's own
    rollout at 1e-6 (step 1e-4: the parameter sensitivities are ~1e-4 of L)."""
    ns = jittered_lattice(16, seed=4, h_over_s=1.8)
    ic = _ic(ns)
    U0 = _state(ic)
    steps, every = 25, 7
    print("[adjoint] rollout parameter gradient", flush=True)
    s = _device(ns, PH)
    s.set_state(U0, dt=DT)
    s.adj_setup(every)
    s.rollout_forward(steps)
    s.rollout_vjp(1.0)
    pg = s.adj_param_grad()
    zero = {k: np.zeros(ns.npfb) for k in KEYS}
    h = 1e-4
    for name in ("Wi", "beta", "eps"):
        _, dLo, _ = _oracle(ns, PH, ic).rollout_tlm(zero, steps, {name: 1.0})
        assert abs(pg[name] - dLo) <= 1e-9 * abs(dLo), (name, pg[name], dLo)
        Lpm = []
        for sign in (1, -1):
            s.set_phys(**dict(PH, **{name: PH[name] + sign * h}))
            s.set_state(U0, dt=DT)
            Lpm.append(s.rollout_forward(steps))
        s.set_phys(**PH)
        fd = (Lpm[0] - Lpm[1]) / (2 * h)
        assert abs(fd - pg[name]) <= 1e-6 * abs(pg[name]), (name, fd, pg[name])
    s.close()


def test_torch_autograd_parameters():
    torch = pytest.importorskip("torch")
    from sunset_amd.autograd import rollout_loss
    ns = jittered_lattice(16, seed=6, h_over_s=1.8)
    ic = _ic(ns)
    U0 = torch.tensor(np.stack([_state(ic)[k] for k in KEYS]), dtype=torch.float64)
    s = _device(ns, PH)
    Wi = torch.tensor(PH["Wi"], dtype=torch.float64, requires_grad=True)
    beta = torch.tensor(PH["beta"], dtype=torch.float64, requires_grad=True)
    eps_ = torch.tensor(PH["eps"], dtype=torch.float64, requires_grad=True)
    L = rollout_loss(s, U0, 12, DT, checkpoint_every=5, Wi=Wi, beta=beta, eps=eps_)
    L.backward()
    h = 1e-4
    with torch.no_grad():
        for name, t in (("Wi", Wi), ("beta", beta), ("eps", eps_)):
            def at(sign):
                kw = {k: torch.tensor(PH[k] + (sign * h if k == name else 0.0),
                                      dtype=torch.float64) for k in ("Wi", "beta", "eps")}
                return float(rollout_loss(s, U0, 12, DT, 5, **kw))
            Lp, Lm = at(+1), at(-1)
            fd = (Lp - Lm) / (2 * h)
            assert abs(fd - float(t.grad)) <= 1e-6 * abs(fd), (name, fd, float(t.grad))
    s.close()


def test_torch_autograd_device_tensors():
    """CUDA tensors at the boundary (sv_set_state_device, sv_rollout_vjp_device)
    in a fresh process that imports torch before the library, so the two
    share one HIP runtime; bit-for-bit equal to the CPU-tensor path
    (tests/gpu_torch_device_check.py)."""
    import os
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    print("[gpu-adjoint] device tensors (subprocess)", flush=True)
    r = subprocess.run([sys.executable, "-u", os.path.join(here, "gpu_torch_device_check.py")],
                       capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0 and r.stdout.strip().endswith("OK")

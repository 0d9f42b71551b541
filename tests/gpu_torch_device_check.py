"""Device-tensor binding check, run as its own process by
tests/test_gpu_adjoint.py::test_torch_autograd_device_tensors: torch is
imported before the library, so the two share one HIP runtime and a CUDA
tensor's data pointer is device memory of the library's runtime.

Checks: one HIP runtime mapped; the state set from CUDA tensors reads back
bit for bit; the rollout loss and its gradient with U0 a CUDA tensor (no host
copy at the boundary) equal the CPU-tensor path bit for bit, and the gradient
is a CUDA tensor; Wi, beta, eps gradients equal too; a host pointer is
refused.  Prints one line per check and "OK" at the end."""
import os
import sys

import torch  # first: the library then binds to torch's HIP runtime

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "sunset-viscoelastic_amd"))

import sunset_amd  # noqa: E402
from sunset_amd.autograd import rollout_loss  # noqa: E402
from sunset_amd.nodes import jittered_lattice, perturbed_rest_ic  # noqa: E402

KEYS = ("ro", "rou", "rov", "cxx", "cxy", "cyy")
PH = dict(Re=1.0, Wi=1.0, beta=0.5, eps=0.1, Ma=0.1, kol_amp=1.0, kol_k=4 * np.pi)
DT = 2e-4


def main():
    assert torch.cuda.is_available(), "no GPU visible to torch"
    runtimes = {ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64" in ln}
    print("HIP runtimes mapped:", sorted(runtimes), flush=True)
    assert len(runtimes) == 1
    ns = jittered_lattice(16, seed=6, h_over_s=1.8)
    ic = perturbed_rest_ic(ns, amp=0.05, seed=3)
    ic["cxx"] = 1.0 + 0.2 * np.sin(2 * np.pi * ns.x) ** 2
    ic["cxy"] = 0.1 * np.cos(2 * np.pi * ns.y)
    U0 = np.stack([ic["ro"], ic["ro"] * ic["u"], ic["ro"] * ic["v"], ic["cxx"], ic["cxy"],
                   ic["cyy"]])
    dev = torch.device("cuda:0")

    def solver():
        s = sunset_amd.Solver(morder=4, form="di").setup(ns)
        s.set_phys(**PH)
        return s

    # set from device tensors, read back into device tensors
    s = solver()
    s.set_state({k: U0[q] for q, k in enumerate(KEYS)}, dt=DT)
    Ud = torch.tensor(U0 * 1.01, dtype=torch.float64, device=dev)
    s.set_state_device([Ud[q].data_ptr() for q in range(6)],
                       torch.cuda.current_stream(dev).cuda_stream)
    back = torch.zeros_like(Ud)
    s.get_state_device([back[q].data_ptr() for q in range(6)])
    assert torch.equal(back, Ud), "device state round trip"
    print("device state round trip: bitwise", flush=True)
    # a host pointer is refused (no device access happens)
    host = torch.zeros(6, ns.npfb, dtype=torch.float64)
    try:
        s.set_state_device([host[q].data_ptr() for q in range(6)])
        raise AssertionError("host pointer accepted")
    except sunset_amd.SunsetError as e:
        assert "not device memory" in str(e)
    print("host pointer refused", flush=True)
    s.close()

    def run(U, params):
        s = solver()
        x = U.clone().requires_grad_(True)
        ps = {k: torch.tensor(PH[k], dtype=torch.float64, requires_grad=True)
              for k in params}
        L = rollout_loss(s, x, 12, DT, checkpoint_every=5, **ps)
        L.backward()
        s.close()
        return L.detach(), x.grad, {k: float(v.grad) for k, v in ps.items()}

    Lh, Gh, Ph = run(torch.tensor(U0, dtype=torch.float64), ("Wi", "beta", "eps"))
    Ld, Gd, Pd = run(torch.tensor(U0, dtype=torch.float64, device=dev), ("Wi", "beta", "eps"))
    assert Gd.is_cuda and Ld.is_cuda
    assert float(Lh) == float(Ld), (float(Lh), float(Ld))
    assert torch.equal(Gh, Gd.cpu()), float((Gh - Gd.cpu()).abs().max())
    assert Ph == Pd, (Ph, Pd)
    print(f"CUDA-tensor rollout = CPU-tensor rollout bit for bit: loss {float(Ld):.6e}, "
          f"|grad| {float(Gd.norm()):.6e}, dL/deps {Pd['eps']:.6e}", flush=True)
    print("OK", flush=True)


if __name__ == "__main__":
    main()

"""Device-tensor binding check, run as its own process by
tests/test_gpu_adjoint.py::test_torch_autograd_device_tensors: torch is
imported before the library, so the two share one HIP runtime and a CUDA
tensor's data pointer is device memory of the library's runtime.

Checks: one HIP runtime mapped; the state set from CUDA tensors reads back
bit for bit; the rollout loss and its gradient with U0 a CUDA tensor (no host
copy at the boundary) equal the CPU-tensor path bit for bit, and the gradient
is a CUDA tensor; Wi, beta, eps gradients equal too; a host pointer is
refused; the device-field operators (sv_apply_dev on CUDA tensors of the real
nodes, ghosts resolved to their parents) equal the host drop-in operators bit
for bit on a channel with walls, calc_laplacian_transverse_only_on_bound
included.  Prints one line per check and "OK" at the end."""
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
    apply_dev_check(dev)
    print("OK", flush=True)


def apply_dev_check(dev):
    import ctypes as C
    from sunset_amd.nodes import wall_channel
    ns = wall_channel(16, 16, order=6, seed=1)
    s = sunset_amd.Solver(morder=6, form="lc").setup(ns)
    f = np.sin(2 * np.pi * ns.x) * np.cos(np.pi * ns.y) + ns.y ** 3
    _, _, irel, _ = s.ghosts()
    full = np.concatenate([f, f[irel]])   # periodic ghosts: their parents' values
    host = {1: s.laplacian(full), 2: s.gradient_only(1, full), 3: s.gradient_only(2, full),
            6: s.laplacian_transverse(full)}
    L = sunset_amd.lib()
    fd = torch.tensor(f, dtype=torch.float64, device=dev)
    st = torch.cuda.current_stream(dev)
    for op, want in host.items():
        out = torch.full((ns.npfb,), float("nan"), dtype=torch.float64, device=dev)
        pin = (C.c_void_p * 1)(fd.data_ptr())
        pout = (C.c_void_p * 1)(out.data_ptr())
        rc = L.sv_apply_dev(s.h, op, 1, pin, pout, C.c_void_p(st.cuda_stream))
        assert rc == 0, sunset_amd.lib().sv_last_error(s.h)
        got = out.cpu().numpy()   # ordered after the launches on torch's stream
        assert np.array_equal(got, want), (op, float(np.abs(got - want).max()))
    s.close()
    print("device-field operators = host drop-in operators bit for bit (wall channel, "
          "transverse Laplacian included)", flush=True)


if __name__ == "__main__":
    main()

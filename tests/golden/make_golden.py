"""Write the golden fixtures under tests/golden/ from the flang-built reference.

Run in the build container (needs oracle/_ref, built by ``make -C oracle ref``
from the reference sources, and the reference's IPART_nodecomp data file):

    python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):
  ipart_nodecomp.npz   the reference's own node-set data file (4056 nodes,
                       periodic [-pi,pi]^2) as arrays
  ref_<name>.npz       reference outputs on a seeded jittered lattice: ghost
                       positions/parents, neighbour lists (CSR, canonical
                       (r^2, j) order), weights and sums in the same order,
                       derivative operators of the initial density, and the
                       fields after `nsteps` RK3 steps of the log-conformation
                       solver
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "sunset-viscoelastic_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from sunset_amd.nodes import jittered_lattice, read_ipart, taylor_green_ic  # noqa: E402
import refdrv  # noqa: E402

CASES = [
    # name, order, n, h/s, seed, nsteps, physics
    # h = 1.3 s as specified for the 64^2 order-4 case: unstable in the reference
    # itself after ~15 steps, so only the first steps are a parity vector
    ("c1_m4", 4, 32, 1.3, 101, 8, dict(Re=1.0, Wi=1.0, beta=0.5, eps=0.0, Ma=0.1)),
    # the reference's own order-4 stencil size h = 1.8 s (common_parameter.F90:46)
    ("m4_h18", 4, 20, 1.8, 303, 100, dict(Re=1.0, Wi=1.0, beta=0.5, eps=0.0, Ma=0.1)),
    ("m6_20", 6, 20, 2.4, 202, 20, dict(Re=1.0, Wi=2.0, beta=0.1, eps=0.25, Ma=0.1)),
]


def to_csr(out, n):
    """Reference lists -> CSR in canonical (r^2, j) order, weights permuted alike."""
    x, y = out["x"], out["y"]
    cnt = out["ij_count"]
    ptr = np.zeros(n + 1, dtype=np.int64)
    ptr[1:] = np.cumsum(cnt)
    col = np.zeros(ptr[-1], dtype=np.int32)
    wx = np.zeros(ptr[-1]); wy = np.zeros(ptr[-1]); wl = np.zeros(ptr[-1]); wh = np.zeros(ptr[-1])
    for i in range(n):
        js = out["ij_link"][i, :cnt[i]]
        dx = x[i] - x[js]
        dy = y[i] - y[js]
        o = np.lexsort((js, dx * dx + dy * dy))
        s = slice(ptr[i], ptr[i + 1])
        col[s] = js[o]
        wx[s] = out["w_grad"][i, :cnt[i], 0][o]
        wy[s] = out["w_grad"][i, :cnt[i], 1][o]
        wl[s] = out["w_lap"][i, :cnt[i]][o]
        wh[s] = out["w_hyp"][i, :cnt[i]][o]
    return ptr, col, wx, wy, wl, wh


def main():
    ref_ipart = "/root/reference/IPART_nodecomp"
    if os.path.exists(ref_ipart):
        ns = read_ipart(ref_ipart, order=4)
        np.savez_compressed(os.path.join(HERE, "ipart_nodecomp.npz"), x=ns.x, y=ns.y, s=ns.s,
                            node_type=ns.node_type,
                            box=np.array([ns.xmin, ns.xmax, ns.ymin, ns.ymax]),
                            bc=np.array(ns.bc), smax=np.array(ns.smax))
    for name, order, n, hos, seed, nsteps, phys in CASES:
        ns = jittered_lattice(n, seed=seed, h_over_s=hos)
        ic = taylor_green_ic(ns)
        out, _ = refdrv.run_reference(ns, order, ic, nsteps=nsteps, phys=phys)
        ptr, col, wx, wy, wl, wh = to_csr(out, ns.npfb)
        np.savez_compressed(
            os.path.join(HERE, f"ref_{name}.npz"),
            order=order, n=n, h_over_s=hos, seed=seed, nsteps=nsteps,
            phys=np.array([phys["Re"], phys["Wi"], phys["beta"], phys["eps"], phys["Ma"]]),
            x=ns.x, y=ns.y, gx=out["x"][ns.npfb:], gy=out["y"][ns.npfb:], irel=out["irel"],
            ij_count=out["ij_count"], ij_count_small=out["ij_count_small"], ptr=ptr, col=col,
            wx=wx, wy=wy, wl=wl, wh=wh, gsum=out["gsum"], lsum=out["lsum"], hsum=out["hsum"],
            vol=out["vol"], grad_ro_x=out["grad_ro_x"], grad_ro_y=out["grad_ro_y"],
            lap_ro=out["lap_ro"], filt_ro=out["filt_ro"],
            **{f"f_{k}": out[k] for k in ("ro", "rou", "rov", "u", "v", "pxx", "pxy", "pyy",
                                         "cxx", "cxy", "cyy")},
            ctl=np.array([out[k] for k in ("dt", "time", "emax_np1", "emax_n", "emax_nm1",
                                          "dt_cfl", "dt_par")]),
            itime=out["itime"])
        print("wrote", name, ns.npfb, "nodes", int(ptr[-1]), "links")


if __name__ == "__main__":
    main()

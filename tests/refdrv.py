"""Run the flang-built reference (oracle/_ref/sunset_ref_lc_m{4,6,8}) on a node
set and initial state, and parse its binary output (layout written by
oracle/ref_driver.F90).  Test infrastructure only.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")


def ref_binary(order: int) -> str:
    return os.path.join(REF_DIR, f"sunset_ref_lc_m{order}")


def available(order: int) -> bool:
    return os.path.exists(ref_binary(order))


def ref_binary_mixed(forder: int, morder: int) -> str:
    """The reference built with filter order forder > derivative order morder."""
    if forder == morder:
        return ref_binary(forder)
    return os.path.join(REF_DIR, f"sunset_ref_lc_f{forder}m{morder}")


def dropin_binary(order: int) -> str:
    """The reference's time loop linked with this library's Fortran drop-in
    neighbours / labf / derivatives modules (oracle/Makefile target dropin)."""
    return os.path.join(REF_DIR, f"sunset_dropin_lc_m{order}")


def run_reference(ns, order, ic, *, nsteps=0, dt_mode="pid", phys=None, dt0=1.0e-6,
                  time0=0.0, emax=(1e-5, 1e-5, 1e-5), itime0=0, nplink=0, threads=4,
                  timeout=600, stencil=None, fields_only=False, binary=None, inflow=None,
                  adapt=False, transverse=False):
    """stencil (optional): dict(nplink, ij_count, ij_count_small, ij_link (0-based,
    [npfb, nplink]), w_grad [npfb, nplink, 2], w_lap, w_hyp, gsum [npfb, 2], lsum,
    hsum, vol) loaded instead of the reference's own neighbour/weight setup.
    adapt: run adapt_stencils after the first neighbour search (and the search
    again after it); the output gains adapt_h, adapt_h_small, adapt_count,
    adapt_count_small as adapt_stencils left them.
    transverse: also calc_laplacian_transverse_only_on_bound of ro and cxx
    (lt_ro, lt_cxx) and calc_divergence of (rou, rov) (div_rouv) after the
    operators on ro."""
    phys = dict(L_char=1.0, U_char=1.0, Re=1.0, Wi=1.0, beta=0.5, eps=0.0, Ma=0.1,
                Mdiff=0.0, grav=(0, 0, 0), drive=(0, 0, 0)) | (phys or {})
    n = ns.npfb
    has_b = getattr(ns, "rnx", None) is not None
    flags = ((2 if stencil is not None else 0) | (4 if fields_only else 0) | (8 if has_b else 0)
             | (16 if adapt else 0) | (32 if transverse else 0))
    inflow = dict(inflow_type=0, u_control=0, u_start=0.0, u_end=1.0, ramp_time=1.0) | (
        inflow or {})
    hdr = np.array([n, nsteps, 0 if dt_mode == "pid" else 1, nplink, flags, *ns.bc, itime0, 0,
                    0], dtype=np.int32)
    rh = np.array([ns.xmin, ns.xmax, ns.ymin, ns.ymax, ns.smax, phys["L_char"], phys["U_char"],
                   phys["Re"], phys["Wi"], phys["beta"], phys["eps"], phys["Ma"], phys["Mdiff"],
                   *phys["grav"], *phys["drive"], dt0, time0, *emax], dtype=np.float64)
    with tempfile.TemporaryDirectory(prefix="svref_") as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(hdr.tobytes())
            f.write(rh.tobytes())
            for a in (ns.x, ns.y, ns.s, ns.h, ns.h_small):
                f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
            f.write(np.ascontiguousarray(ns.node_type, dtype=np.int32).tobytes())
            for k in ("ro", "u", "v", "cxx", "cxy", "cyy"):
                f.write(np.ascontiguousarray(ic[k], dtype=np.float64).tobytes())
            if has_b:
                f.write(np.ascontiguousarray(ns.rnx, dtype=np.float64).tobytes())
                f.write(np.ascontiguousarray(ns.rny, dtype=np.float64).tobytes())
                f.write(np.array([inflow["inflow_type"], inflow["u_control"]],
                                 dtype=np.int32).tobytes())
                f.write(np.array([inflow["u_start"], inflow["u_end"], inflow["ramp_time"]],
                                 dtype=np.float64).tobytes())
        args = [binary or ref_binary(order), fin, fout]
        if stencil is not None:
            fst = os.path.join(td, "stencil.bin")
            with open(fst, "wb") as f:
                f.write(np.int32(stencil["nplink"]).tobytes())
                f.write(np.ascontiguousarray(stencil["ij_count"], dtype=np.int32).tobytes())
                f.write(np.ascontiguousarray(stencil["ij_count_small"], dtype=np.int32).tobytes())
                f.write(np.ascontiguousarray(stencil["ij_link"] + 1, dtype=np.int32).tobytes())
                for k in ("w_grad", "w_lap", "w_hyp", "gsum", "lsum", "hsum", "vol"):
                    f.write(np.ascontiguousarray(stencil[k], dtype=np.float64).tobytes())
            args.append(fst)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
                   OMP_PLACES="cores")
        res = subprocess.run(args, cwd=td, env=env, capture_output=True, text=True,
                             timeout=timeout)
        if res.returncode != 0 or not os.path.exists(fout):
            # a Fortran STOP exits with status 0 and no output file
            raise RuntimeError(f"reference failed rc={res.returncode}\n{res.stdout[-4000:]}\n"
                               f"{res.stderr[-4000:]}")
        raw = open(fout, "rb").read()
    times = {"steps": []}
    for line in res.stdout.splitlines():
        if "REFTIME" in line:
            parts = line.split()
            times[parts[1]] = float(parts[-1])
        if "REFSTEP" in line:
            times["steps"].append(float(line.split()[-1]))
    return _parse(raw, n, nsteps, fields_only, adapt, transverse), times


def _parse(raw, n, nsteps, fields_only=False, adapt=False, transverse=False):
    off = 0

    def take(cnt, dt):
        nonlocal off
        a = np.frombuffer(raw, dtype=dt, count=cnt, offset=off)
        off += a.nbytes
        return a.copy()

    out = {}
    if adapt:
        assert int(take(1, np.int32)[0]) == n
        out["adapt_h"] = take(n, np.float64)
        out["adapt_h_small"] = take(n, np.float64)
        out["adapt_count"] = take(n, np.int32)
        out["adapt_count_small"] = take(n, np.int32)
    if fields_only:
        return _parse_fields(out, take, n)
    np_, nplink = take(2, np.int32)
    out.update(np=int(np_), nplink=int(nplink))
    out["x"] = take(np_, np.float64)
    out["y"] = take(np_, np.float64)
    ng = np_ - n
    out["irel"] = take(ng, np.int32) - 1   # 0-based parents of ghosts
    out["vrel"] = take(ng, np.int32)
    out["ij_count"] = take(n, np.int32)
    out["ij_count_small"] = take(n, np.int32)
    out["ij_link"] = take(n * nplink, np.int32).reshape(n, nplink) - 1
    out["w_grad"] = take(n * nplink * 2, np.float64).reshape(n, nplink, 2)
    out["w_lap"] = take(n * nplink, np.float64).reshape(n, nplink)
    out["w_hyp"] = take(n * nplink, np.float64).reshape(n, nplink)
    out["gsum"] = take(2 * n, np.float64).reshape(n, 2)
    out["lsum"] = take(n, np.float64)
    out["hsum"] = take(n, np.float64)
    out["vol"] = take(n, np.float64)
    out["grad_ro_x"] = take(n, np.float64)
    out["grad_ro_y"] = take(n, np.float64)
    out["lap_ro"] = take(n, np.float64)
    out["filt_ro"] = take(n, np.float64)
    if transverse:
        out["lt_ro"] = take(n, np.float64)
        out["lt_cxx"] = take(n, np.float64)
        out["div_rouv"] = take(n, np.float64)
    return _parse_fields(out, take, n)


def _parse_fields(out, take, n):
    for k in ("ro", "rou", "rov", "u", "v", "pxx", "pxy", "pyy", "cxx", "cxy", "cyy"):
        out[k] = take(n, np.float64)
    ctl = take(7, np.float64)
    for k, v in zip(("dt", "time", "emax_np1", "emax_n", "emax_nm1", "dt_cfl", "dt_par"), ctl):
        out[k] = float(v)
    out["itime"] = int(take(1, np.int32)[0])
    return out


def canonical_rows(link, count, x, y, xi, yi):
    """Sort each neighbour row by (r^2, j) -- the order the reference's unstable
    quicksort leaves undetermined for ties (neighbours.F90:316-341)."""
    n = count.shape[0]
    out = np.full_like(link, -1)
    for i in range(n):
        js = link[i, :count[i]]
        dx = xi[i] - x[js]
        dy = yi[i] - y[js]
        r2 = dx * dx + dy * dy
        o = np.lexsort((js, r2))
        out[i, :count[i]] = js[o]
    return out

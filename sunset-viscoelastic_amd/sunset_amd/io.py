"""The reference's field dump, node dump and restart files.

Layouts (2D, one simulation, file index k = 10000 as the reference's serial
build):

``data_out/fields_10000_<n>`` and ``restart/fields_10000`` (written by
``output_layer``, output.F90:184-320; read back by ``load_restart_file``,
setup_flow.F90:448-552)::

    !! header !!
    npfb
    eflow_nm1 sum_eflow driving_force(1:3)
    emax_np1 emax_n emax_nm1 dt
    !! end header !!
    ro-rho_char u v vorticity Q alpha cxx cxy cyy czz     (one line per node)

``data_out/nodes_10000`` and ``restart/nodes_10000`` (``output_nodes``,
output.F90:421-476; read by setup_domain.F90:206-242)::

    npfb
    global_index x y s h h_small node_type                 (one line per node)

Numbers are written with 17 significant digits (``repr``), so a dump read
back gives the same doubles, and any whitespace-separated reader (Fortran
list-directed input included) reads them.  The vorticity v_x - u_y and the
Q-criterion -u_x^2/2 - v_y^2/2 - u_y v_x come from the device gradient
operator (calc_gradient).  The per-node time-stepping error (``alpha_out``,
step.F90:456) is not kept by the device step and is written as 0; the
restart reader skips that column as the reference's does.  Flow-rate control
(eflow_nm1, sum_eflow) is not on this path: its restart values are written
as the reference's initial ones (1, 0, setup_flow.F90:167-168).
"""
from __future__ import annotations

import os

import numpy as np

from .nodes import NodeSet

FILE_INDEX = 10000   # the serial build's processor tag (output.F90:252-256)


def _fmt(v) -> str:
    return repr(float(v))


def _num(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def field_dump_name(n_out: int, out_dir: str = "data_out") -> str:
    return os.path.join(out_dir, f"fields_{FILE_INDEX}_{n_out}")


def _ghost_extend(solver, f, sign_u: bool = False, sign_v: bool = False):
    """Values on real + ghost nodes: ghosts copy their parent, with the sign
    flips of the mirror relation for velocity components
    (mirror_boundaries.F90:263-352)."""
    _, _, irel, vrel = solver.ghosts()
    g = f[irel].copy()
    if sign_u:
        g[(vrel == 2) | (vrel == 4)] *= -1.0
    if sign_v:
        g[(vrel == 3) | (vrel == 4)] *= -1.0
    return np.concatenate([f, g])


def write_fields(path: str, solver, rho_char: float = 1.0, drive=(0.0, 0.0, 0.0),
                 eflow_nm1: float = 1.0, sum_eflow: float = 0.0) -> None:
    """Dump the solver's current state in the output_layer layout."""
    F = solver.fields()
    u, v = F["u"], F["v"]
    ux, uy = solver.gradient(_ghost_extend(solver, u, sign_u=True))
    vx, vy = solver.gradient(_ghost_extend(solver, v, sign_v=True))
    vort = vx - uy
    qcrit = -0.5 * ux * ux - 0.5 * vy * vy - uy * vx
    n = solver.npfb
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    cols = np.column_stack([F["ro"] - rho_char, u, v, vort, qcrit, np.zeros(n), F["cxx"],
                            F["cxy"], F["cyy"], np.ones(n)])
    with open(path, "w") as f:
        f.write(" !! header !!\n")
        f.write(f" {n}\n")
        f.write(" " + " ".join(_fmt(x) for x in (eflow_nm1, sum_eflow, *drive)) + "\n")
        f.write(" " + " ".join(_fmt(F[k]) for k in ("emax_np1", "emax_n", "emax_nm1", "dt"))
                + "\n")
        f.write(" !! end header !!\n")
        for row in cols:
            f.write(" " + " ".join(_fmt(x) for x in row) + "\n")


def read_fields(path: str, rho_char: float = 1.0) -> dict:
    """Read a field dump as load_restart_file does: the initial condition
    (ro, u, v, cxx, cxy, cyy, czz) and the PID controller state."""
    with open(path) as f:
        lines = f.read().splitlines()
    npfb = int(lines[1].split()[0])
    h1 = [_num(x) for x in lines[2].split()]
    h2 = [_num(x) for x in lines[3].split()]
    body = np.array([[_num(x) for x in ln.split()] for ln in lines[5:5 + npfb]])
    if body.shape != (npfb, 10):
        raise ValueError(f"{path}: expected {npfb} rows of 10 values, got {body.shape}")
    return dict(npfb=npfb, ro=body[:, 0] + rho_char, u=body[:, 1], v=body[:, 2],
                cxx=body[:, 6], cxy=body[:, 7], cyy=body[:, 8], czz=body[:, 9],
                eflow_nm1=h1[0], sum_eflow=h1[1], driving_force=tuple(h1[2:5]),
                emax_np1=h2[0], emax_n=h2[1], emax_nm1=h2[2], dt=h2[3])


def write_nodes(path: str, ns: NodeSet) -> None:
    """output_nodes layout: count, then gid x y s h h_small node_type."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "w") as f:
        f.write(f" {ns.npfb}\n")
        for i in range(ns.npfb):
            f.write(f" {i + 1} {_fmt(ns.x[i])} {_fmt(ns.y[i])} {_fmt(ns.s[i])} "
                    f"{_fmt(ns.h[i])} {_fmt(ns.h_small[i])} {int(ns.node_type[i])}\n")


def read_nodes_restart(path: str, ns: NodeSet) -> NodeSet:
    """Apply a nodes restart file (setup_domain.F90:206-242): h and h_small
    per node, checking the global index and node type of every line."""
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    n = int(lines[0][0])
    if n != ns.npfb:
        raise ValueError(f"{path}: {n} nodes, node set has {ns.npfb}")
    h = np.empty(n)
    hs = np.empty(n)
    for i, r in enumerate(lines[1:1 + n]):
        if int(r[0]) != i + 1:
            raise ValueError(f"{path}: global index mismatch at line {i + 2}: {r[0]}")
        if int(r[6]) != int(ns.node_type[i]):
            raise ValueError(f"{path}: node type mismatch at node {i + 1}")
        h[i], hs[i] = _num(r[4]), _num(r[5])
    out = NodeSet(**{f.name: getattr(ns, f.name) for f in ns.__dataclass_fields__.values()})
    out.h, out.h_small = h, hs
    return out


def restart(solver, path: str, time: float = 0.0, itime: int = 0, rho_char: float = 1.0):
    """Set the solver's state and PID controller from a field dump (the
    solver's nodes, weights and physics already set up)."""
    R = read_fields(path, rho_char)
    if R["npfb"] != solver.npfb:
        raise ValueError(f"{path}: {R['npfb']} nodes, solver has {solver.npfb}")
    solver.set_initial(R, dt=R["dt"], itime=itime, time=time,
                       emax=(R["emax_np1"], R["emax_n"], R["emax_nm1"]))
    return R

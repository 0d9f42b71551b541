"""Node sets, the IPART text format and the benchmark configurations.

The IPART format is the reference's node-set file (read by
``setup_domain.F90:93-199``, written by ``source/gen/datclass.F90:734-745``)::

    nb npfb smax smin
    xmin xmax ymin ymax
    xbcL xbcU ybcL ybcU            (0 none, 1 periodic, 2 symmetric, 3 no-slip)
    [nprocsX nprocsY nprocsZ / nprocs / nprocs lines "start end"]   (optional)
    gid x y type nx ny s           (one line per node; type 999 = fluid)

Node sets here are seeded jittered lattices (numpy PCG64, seed recorded in
``NodeSet.meta``), written in x-major order like the reference generator's
x-sorted output (``datclass.F90:718``).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

# stencil-size multipliers h = hovs*s per order, common_parameter.F90:45-56
HOVS = {4: 1.8, 6: 2.4, 8: 2.7, 10: 3.1}
# h = hovs_bound*s on boundary nodes and their FD rows (setup_domain.F90:182)
HOVS_BOUND = {4: 2.1, 6: 2.4, 8: 2.4, 10: 2.4}


@dataclasses.dataclass
class NodeSet:
    x: np.ndarray
    y: np.ndarray
    s: np.ndarray
    h: np.ndarray
    h_small: np.ndarray
    node_type: np.ndarray
    xmin: float
    xmax: float
    ymin: float
    ymax: float
    bc: tuple = (1, 1, 1, 1)
    smax: float = 0.0
    meta: dict = dataclasses.field(default_factory=dict)
    # unit normals into the fluid of boundary nodes (types 0 wall, 1 inflow,
    # 2 outflow) and their FD rows (types -1..-4, which follow their boundary
    # node in index order); None when the set has no boundary nodes
    rnx: Optional[np.ndarray] = None
    rny: Optional[np.ndarray] = None

    @property
    def nb(self) -> int:
        return int(np.count_nonzero((self.node_type >= 0) & (self.node_type <= 2)))

    @property
    def npfb(self) -> int:
        return int(self.x.shape[0])


def jittered_lattice(nx: int, ny: Optional[int] = None, *, seed: int = 0,
                     jitter: float = 0.3, h_over_s: float = 1.3,
                     hs_over_s: Optional[float] = None, box=(0.0, 1.0, 0.0, 1.0),
                     bc=(1, 1, 1, 1)) -> NodeSet:
    """Periodic jittered lattice: x = (i+0.5)s + U(-jitter, jitter)s."""
    ny = nx if ny is None else ny
    x0, x1, y0, y1 = box
    sx = (x1 - x0) / nx
    sy = (y1 - y0) / ny
    s = min(sx, sy)
    rng = np.random.Generator(np.random.PCG64(seed))
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ii = ii.ravel()
    jj = jj.ravel()
    jx = rng.uniform(-jitter, jitter, ii.shape[0])
    jy = rng.uniform(-jitter, jitter, ii.shape[0])
    x = x0 + (ii + 0.5) * sx + jx * sx
    y = y0 + (jj + 0.5) * sy + jy * sy
    n = x.shape[0]
    sv = np.full(n, s)
    h = sv * h_over_s
    hs = sv * (h_over_s if hs_over_s is None else hs_over_s)
    return NodeSet(x=x, y=y, s=sv, h=h, h_small=hs,
                   node_type=np.full(n, 999, dtype=np.int32),
                   xmin=x0, xmax=x1, ymin=y0, ymax=y1, bc=tuple(bc), smax=s,
                   meta=dict(generator="PCG64", seed=seed, nx=nx, ny=ny,
                             jitter=jitter, h_over_s=h_over_s))


def write_ipart(path: str, ns: NodeSet, nprocs_schedule: bool = False) -> None:
    """Write the IPART text file (datclass.F90:734-745 layout).  FD rows
    (types -1..-4) are not written: the reader generates them from their
    boundary node's position, normal and spacing (setup_domain.F90:179-193)."""
    nb = ns.nb
    keep = np.where(ns.node_type >= 0)[0]
    nx = ns.rnx if ns.rnx is not None else np.zeros(ns.npfb)
    ny = ns.rny if ns.rny is not None else np.zeros(ns.npfb)
    with open(path, "w") as f:
        f.write(f"{nb} {keep.shape[0]} {float(ns.smax)!r} {float(ns.s.min())!r}\n")
        f.write(f"{float(ns.xmin)!r} {float(ns.xmax)!r} {float(ns.ymin)!r} "
                f"{float(ns.ymax)!r}\n")
        f.write(" ".join(str(int(b)) for b in ns.bc) + "\n")
        if nprocs_schedule:
            f.write("1 1 1\n1\n")
            f.write(f"1 {keep.shape[0]}\n")
        for g, i in enumerate(keep):
            f.write(f"{g + 1} {float(ns.x[i])!r} {float(ns.y[i])!r} {int(ns.node_type[i])} "
                    f"{float(nx[i])!r} {float(ny[i])!r} {float(ns.s[i])!r}\n")


def read_ipart(path: str, order: int = 8) -> NodeSet:
    """Read an IPART file (with or without a decomposition schedule).

    h = hovs*s as setup_domain.F90:180 (hovs for the given order), h_small = h.
    A boundary node (type 0 wall, 1 inflow, 2 outflow) gets h = hovs_bound*s
    and is followed by its four FD rows at s, 2s, 3s, 4s along its normal
    (types -1..-4, same normal, h and s), as setup_domain.F90:179-193.
    """
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    nb, npfb = int(lines[0][0]), int(lines[0][1])
    smax = float(lines[0][2].replace("D", "E").replace("d", "e"))
    xmin, xmax, ymin, ymax = (float(v) for v in lines[1][:4])
    bc = tuple(int(v) for v in lines[2][:4])
    k = 3
    # optional schedule: a 3-int line, then nprocs, then nprocs pairs
    if len(lines[k]) == 3 and all(_isint(v) for v in lines[k]):
        nprocs = int(lines[k + 1][0])
        k += 2 + nprocs
    rows = lines[k:k + npfb]
    arr = np.array([[float(v.replace("D", "E").replace("d", "e")) for v in r[:7]]
                    for r in rows])
    X, Y, T, S, H, NX, NY = [], [], [], [], [], [], []
    for r in arr:
        t = int(r[3])
        X.append(r[1]); Y.append(r[2]); T.append(t); S.append(r[6]); NX.append(r[4]); NY.append(r[5])
        if 0 <= t <= 2:
            hb = r[6] * HOVS_BOUND.get(order, 2.4)
            H.append(hb)
            for j in range(1, 5):
                X.append(r[1] + r[4] * float(j) * r[6]); Y.append(r[2] + r[5] * float(j) * r[6])
                T.append(-j); S.append(r[6]); H.append(hb); NX.append(r[4]); NY.append(r[5])
        else:
            H.append(r[6] * HOVS.get(order, 2.7))
    typ = np.array(T, dtype=np.int32)
    h = np.array(H)
    has_b = bool(np.any(typ != 999) and np.any((typ >= 0) & (typ <= 2)))
    return NodeSet(x=np.array(X), y=np.array(Y), s=np.array(S), h=h, h_small=h.copy(),
                   node_type=typ, xmin=xmin, xmax=xmax, ymin=ymin, ymax=ymax, bc=bc, smax=smax,
                   meta=dict(source=path, nb=nb),
                   rnx=np.array(NX) if has_b else None, rny=np.array(NY) if has_b else None)


def _isint(v: str) -> bool:
    try:
        int(v)
        return True
    except ValueError:
        return False


def nplink_default(order: int) -> int:
    """initial_setup's neighbour-list capacity, setup_domain.F90:50."""
    return 2 * 4 * int(math.ceil(2.0 * HOVS.get(order, 2.7))) ** 2


def taylor_green_ic(ns: NodeSet):
    """u = sin(2 pi x)cos(2 pi y), v = -cos(2 pi x) sin(2 pi y), rho = 1, c = I."""
    n = ns.npfb
    u = np.sin(2 * np.pi * ns.x) * np.cos(2 * np.pi * ns.y)
    v = -np.cos(2 * np.pi * ns.x) * np.sin(2 * np.pi * ns.y)
    return dict(ro=np.ones(n), u=u, v=v, cxx=np.ones(n), cxy=np.zeros(n),
                cyy=np.ones(n))


def perturbed_rest_ic(ns: NodeSet, amp: float = 0.01, seed: int = 0, kmax: int = 4):
    """Fluid at rest (rho = 1, c = I) plus a divergence-free random-phase velocity
    perturbation: streamfunction sum_k a_k sin(2 pi k.x + phi_k) / |k| over
    integer wavevectors 0 < |k_x|, |k_y| <= kmax (a_k ~ U(0,1), phi_k ~ U(0, 2 pi),
    numpy PCG64 with `seed`), u = d(psi)/dy, v = -d(psi)/dx, scaled so that
    max |u, v| over the nodes equals amp."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x, y = ns.x, ns.y
    u = np.zeros_like(x)
    v = np.zeros_like(x)
    for kx in range(-kmax, kmax + 1):
        for ky in range(1, kmax + 1):
            a = rng.uniform(0.0, 1.0)
            ph = rng.uniform(0.0, 2 * np.pi)
            kk = math.hypot(kx, ky)
            arg = 2 * np.pi * (kx * x + ky * y) + ph
            u += a * 2 * np.pi * ky * np.cos(arg) / kk
            v -= a * 2 * np.pi * kx * np.cos(arg) / kk
    scale = amp / max(np.abs(u).max(), np.abs(v).max())
    n = ns.npfb
    return dict(ro=np.ones(n), u=u * scale, v=v * scale, cxx=np.ones(n), cxy=np.zeros(n),
                cyy=np.ones(n))


def ensemble_specs(members: int = 32, seed: int = 7):
    """Members of the batched-ensemble workload: (lattice side, LABFM order,
    Wi, beta) with node counts N = round(exp(U(ln 16384, ln 262144))) (side =
    round(sqrt N), at least 16), order 4 or 6, Wi = exp(U(ln 0.1, ln 10)) and
    beta = U(0.1, 0.9), drawn in that order per member from numpy PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    specs = []
    for _ in range(members):
        N = int(round(math.exp(rng.uniform(math.log(16384), math.log(262144)))))
        side = max(16, int(round(math.sqrt(N))))
        order = int(rng.choice([4, 6]))
        Wi = math.exp(rng.uniform(math.log(0.1), math.log(10.0)))
        beta = rng.uniform(0.1, 0.9)
        specs.append((side, order, Wi, beta))
    return specs


def cylinder_channel(smin: float = 0.004, smax: float = 0.016, grow: float = 1.5,
                     box=(0.0, 8.0, 0.0, 2.0), centre=(2.0, 1.0), radius: float = 0.5,
                     seed: int = 3, jitter: float = 0.3, h_over_s: float = 2.7,
                     h_over_s_wall: float = 2.2, wall_band: float = 3.0):
    """Channel with a cylinder, spacing growing from smin at the cylinder to smax
    (linearly over `grow`), advancing-front placement (Fornberg & Flyer 2015),
    seeded jitter of +/- jitter*s, nodes on the cylinder surface at spacing smin.

    Returns (NodeSet, order_map): LABFM order 8 in the interior and 4 within
    `wall_band` local spacings of any boundary (channel walls, inlet, outlet,
    cylinder).  All nodes are interior LABFM nodes (node_type 999); h = h_over_s*s
    for order 8 and h_over_s_wall*s for order 4.  Boundaries are not periodic.
    """
    import ctypes as C
    import os
    lib = C.CDLL(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                              "lib", "libsv_nodegen.so"))
    f = lib.sv_nodegen_cylinder
    f.restype = C.c_int64
    f.argtypes = [C.c_double] * 10 + [C.c_uint64, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.c_int64]
    x0, x1, y0, y1 = box
    cap = int(2.5 * (x1 - x0) * (y1 - y0) / smin ** 2) + 1000
    xs = np.zeros(cap)
    ys = np.zeros(cap)
    n = f(x0, x1, y0, y1, centre[0], centre[1], radius, smin, smax, grow, seed,
          xs.ctypes.data_as(C.POINTER(C.c_double)), ys.ctypes.data_as(C.POINTER(C.c_double)),
          cap)
    if n < 0 or n > cap:
        raise RuntimeError("node generation failed")
    xs, ys = xs[:n], ys[:n]

    def spacing(x, y):
        d = np.maximum(np.hypot(x - centre[0], y - centre[1]) - radius, 0.0)
        return smin + (smax - smin) * np.minimum(d / grow, 1.0)

    s = spacing(xs, ys)
    d_cyl = np.hypot(xs - centre[0], ys - centre[1]) - radius
    margin = 0.5 * s
    keep = ((d_cyl > margin) & (xs > x0 + margin) & (xs < x1 - margin) & (ys > y0 + margin)
            & (ys < y1 - margin))
    xs, ys, s = xs[keep], ys[keep], s[keep]
    rng = np.random.Generator(np.random.PCG64(seed))
    xs = xs + rng.uniform(-jitter, jitter, xs.shape) * s
    ys = ys + rng.uniform(-jitter, jitter, ys.shape) * s
    # nodes on the cylinder surface
    nc = int(round(2 * np.pi * radius / smin))
    th = np.arange(nc) * 2 * np.pi / nc
    xc = centre[0] + radius * np.cos(th)
    yc = centre[1] + radius * np.sin(th)
    x = np.concatenate([xs, xc])
    y = np.concatenate([ys, yc])
    s = np.concatenate([s, np.full(nc, smin)])
    # x-major order, like the reference's generator output (datclass.F90:718)
    o = np.argsort(x, kind="stable")
    x, y, s = x[o], y[o], s[o]
    d_wall = np.minimum.reduce([np.hypot(x - centre[0], y - centre[1]) - radius, x - x0,
                                x1 - x, y - y0, y1 - y])
    order = np.where(d_wall <= wall_band * s, 4, 8).astype(np.int32)
    h = np.where(order == 8, h_over_s, h_over_s_wall) * s
    ns = NodeSet(x=x, y=y, s=s, h=h, h_small=h.copy(),
                 node_type=np.full(x.shape[0], 999, dtype=np.int32), xmin=x0, xmax=x1,
                 ymin=y0, ymax=y1, bc=(0, 0, 0, 0), smax=smax,
                 meta=dict(generator="advancing front (Fornberg-Flyer 2015) + PCG64 jitter",
                           seed=seed, smin=smin, smax=smax, grow=grow))
    return ns, order


def _fd_column(xb, yb, nx_, ny_, s, hb, tb):
    """A boundary node and its four FD rows along the normal
    (setup_domain.F90:179-193)."""
    xs = [xb + nx_ * j * s for j in range(5)]
    ys = [yb + ny_ * j * s for j in range(5)]
    return xs, ys, [tb, -1, -2, -3, -4]


def wall_channel(nx: int, ny: int, *, order: int = 6, seed: int = 0, jitter: float = 0.3,
                 box=(0.0, 1.0, 0.0, 1.0)) -> NodeSet:
    """Channel periodic in x with no-slip walls (node type 0) at y = ymin and
    y = ymax.  Spacing s = (xmax-xmin)/nx; ny = (ymax-ymin)/s layers.  Wall
    nodes sit at x = (i+0.5)s with their FD rows at 1..4 s along the normal;
    the interior is a jittered lattice on the layers 5 .. ny-5 (PCG64, seed).
    Nodes are written column by column, a boundary node followed by its rows.
    h = hovs*s inside, hovs_bound*s on walls and rows (setup_domain.F90:180-190)."""
    x0, x1, y0, y1 = box
    s = (x1 - x0) / nx
    if abs((y1 - y0) / s - ny) > 1e-9 or ny < 12:
        raise ValueError("ny must equal (ymax-ymin)/s and be >= 12")
    rng = np.random.Generator(np.random.PCG64(seed))
    X, Y, T, NX, NY = [], [], [], [], []
    for i in range(nx):
        xc = x0 + (i + 0.5) * s
        xs, ys, ts = _fd_column(xc, y0, 0.0, 1.0, s, 0, 0)
        X += xs; Y += ys; T += ts; NX += [0.0] * 5; NY += [1.0] * 5
        for m in range(5, ny - 4):
            X.append(xc + rng.uniform(-jitter, jitter) * s)
            Y.append(y0 + m * s + rng.uniform(-jitter, jitter) * s)
            T.append(999); NX.append(0.0); NY.append(0.0)
        xs, ys, ts = _fd_column(xc, y1, 0.0, -1.0, s, 0, 0)
        X += xs; Y += ys; T += ts; NX += [0.0] * 5; NY += [-1.0] * 5
    x, y = np.array(X), np.array(Y)
    t = np.array(T, dtype=np.int32)
    sv = np.full(x.shape[0], s)
    h = np.where(t == 999, HOVS[order], HOVS_BOUND[order]) * s
    return NodeSet(x=x, y=y, s=sv, h=h, h_small=h.copy(), node_type=t, xmin=x0, xmax=x1,
                   ymin=y0, ymax=y1, bc=(1, 1, 0, 0), smax=s,
                   meta=dict(generator="PCG64", seed=seed, nx=nx, ny=ny, jitter=jitter,
                             kind="wall_channel"),
                   rnx=np.array(NX), rny=np.array(NY))


def inflow_outflow_channel(nx: int, ny: int, *, order: int = 6, seed: int = 0,
                           jitter: float = 0.3, box=(0.0, 1.0, 0.0, 1.0)) -> NodeSet:
    """Channel periodic in y with an inflow (node type 1, normal +x) at
    x = xmin and an outflow (type 2, normal -x) at x = xmax; spacing
    s = (ymax-ymin)/ny, nx = (xmax-xmin)/s layers; layout as wall_channel."""
    x0, x1, y0, y1 = box
    s = (y1 - y0) / ny
    if abs((x1 - x0) / s - nx) > 1e-9 or nx < 12:
        raise ValueError("nx must equal (xmax-xmin)/s and be >= 12")
    rng = np.random.Generator(np.random.PCG64(seed))
    X, Y, T, NX, NY = [], [], [], [], []
    for j in range(ny):
        yc = y0 + (j + 0.5) * s
        xs, ys, ts = _fd_column(x0, yc, 1.0, 0.0, s, 1, 1)
        X += xs; Y += ys; T += [1, -1, -2, -3, -4]; NX += [1.0] * 5; NY += [0.0] * 5
    for m in range(5, nx - 4):
        for j in range(ny):
            X.append(x0 + m * s + rng.uniform(-jitter, jitter) * s)
            Y.append(y0 + (j + 0.5) * s + rng.uniform(-jitter, jitter) * s)
            T.append(999); NX.append(0.0); NY.append(0.0)
    for j in range(ny):
        yc = y0 + (j + 0.5) * s
        xs, ys, ts = _fd_column(x1, yc, -1.0, 0.0, s, 2, 2)
        X += xs; Y += ys; T += [2, -1, -2, -3, -4]; NX += [-1.0] * 5; NY += [0.0] * 5
    x, y = np.array(X), np.array(Y)
    t = np.array(T, dtype=np.int32)
    sv = np.full(x.shape[0], s)
    h = np.where(t == 999, HOVS[order], HOVS_BOUND[order]) * s
    return NodeSet(x=x, y=y, s=sv, h=h, h_small=h.copy(), node_type=t, xmin=x0, xmax=x1,
                   ymin=y0, ymax=y1, bc=(0, 0, 1, 1), smax=s,
                   meta=dict(generator="PCG64", seed=seed, nx=nx, ny=ny, jitter=jitter,
                             kind="inflow_outflow_channel"),
                   rnx=np.array(NX), rny=np.array(NY))


def cylinder_in_box(n: int, *, radius: float = 0.2, order: int = 6, seed: int = 0,
                    jitter: float = 0.3, box=(0.0, 1.0, 0.0, 1.0)) -> NodeSet:
    """A periodic box with a no-slip cylinder (curved wall, node type 0) at its
    centre.  Spacing s = (xmax-xmin)/n; wall nodes at angles 2 pi k/m on the
    circle (m = round(2 pi r / s)) with outward normals and their FD rows at
    r + s .. r + 4s; the rest a jittered lattice (PCG64, seed) outside
    r + 4.5 s.  Walls first, each followed by its rows, then the lattice in
    x-major order."""
    x0, x1, y0, y1 = box
    s = (x1 - x0) / n
    cx, cy = 0.5 * (x0 + x1), 0.5 * (y0 + y1)
    m = int(round(2 * np.pi * radius / s))
    X, Y, T, NX, NY = [], [], [], [], []
    for k in range(m):
        th = 2 * np.pi * k / m
        nx_, ny_ = float(np.cos(th)), float(np.sin(th))
        xs, ys, ts = _fd_column(cx + radius * nx_, cy + radius * ny_, nx_, ny_, s, 0, 0)
        X += xs; Y += ys; T += ts; NX += [nx_] * 5; NY += [ny_] * 5
    lat = jittered_lattice(n, seed=seed, jitter=jitter, box=box)
    keep = np.hypot(lat.x - cx, lat.y - cy) > radius + 4.5 * s
    X += list(lat.x[keep]); Y += list(lat.y[keep])
    nk = int(keep.sum())
    T += [999] * nk; NX += [0.0] * nk; NY += [0.0] * nk
    x, y = np.array(X), np.array(Y)
    t = np.array(T, dtype=np.int32)
    sv = np.full(x.shape[0], s)
    h = np.where(t == 999, HOVS[order], HOVS_BOUND[order]) * s
    return NodeSet(x=x, y=y, s=sv, h=h, h_small=h.copy(), node_type=t, xmin=x0, xmax=x1,
                   ymin=y0, ymax=y1, bc=(1, 1, 1, 1), smax=s,
                   meta=dict(generator="PCG64", seed=seed, n=n, radius=radius, jitter=jitter,
                             kind="cylinder_in_box"),
                   rnx=np.array(NX), rny=np.array(NY))

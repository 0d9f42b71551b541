#!/usr/bin/env python3
"""Benchmark: RK3 node-steps/s in FP64 on the headline workload.

Workload (one MI355X): 1024 x 1024 seeded jittered periodic lattice (~1M
nodes) on the unit square, LABFM order 6, h = h_small = 2.4 s, log-conformation
sPTT (eps = 0.25, Wi = 2, beta = 0.1, Re = 1, Ma = 0.1), Kolmogorov forcing
f_x = sin(4 pi y), PID time step from dt = 1e-6 as the reference starts.  A
step is one full RK3(2)4[2R+]C step: 4 right-hand sides + stage updates,
filter, mass correction, time-step control, all on the device with inputs
resident in HBM.

Printed line (one JSON object): value = node-steps/s (nodes x steps / time of
the timed steps); roofline of the dominant kernel (the RK stage sweep)
measured live with HIP events on the library's stream; weight solves/s of the
device weight setup; cpu_baseline = the reference's own OpenMP RK3 loop
(flang build in oracle/_ref) on the same 1M-node workload (3 timed steps),
on the CPUs this job may use; error_vs_oracle of every leg against the FP64
CPU oracle (and the reference's own fields for the headline run).
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sunset-viscoelastic_amd"))

METRIC = "RK3 node-steps/s FP64 @1M nodes; weight solves/s; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
PHYS = dict(Re=1.0, Wi=2.0, beta=0.1, eps=0.25, Ma=0.1, kol_amp=1.0, kol_k=4.0 * math.pi)


PROFILE_STAGE = os.path.join(ROOT, "profiles", "k_stage_h_lc_sptt_1024sq_order6.txt")


def measured_stage_traffic():
    """HBM bytes per k_stage_h launch (mean of stages 1-4) measured with
    rocprofv3 FETCH_SIZE / WRITE_SIZE on this workload, copy-kernel
    calibrated; the summary and its commands are in the profile file."""
    try:
        with open(PROFILE_STAGE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def stage_bytes_per_node(ks_mean, halo):
    """Algorithmic HBM bytes per node of the four LC stage sweeps (mean).
    Per link: a 16-bit block-local index + 3 FP64 weights (26 B); per staged
    halo record: its 4-byte index + the 9 gathered fields (76 B); own-node
    traffic (stage 1: 316 B, stages 2-3: 364 B, stage 4: 244 B).  Without
    halos (halo == 0): a 4-byte index per link and the gathers are not
    counted."""
    link = 26.0 if halo > 0 else 28.0
    return link * ks_mean + 76.0 * halo + (316.0 + 364.0 + 364.0 + 244.0) / 4.0


def step_bytes_per_node(ks_mean, k_mean, halo):
    """Whole step: 4 stages + filter (10 B per link + 52 B per halo record +
    212 B) + apply (128 B)."""
    flink = 10.0 if halo > 0 else 12.0
    return (4 * stage_bytes_per_node(ks_mean, halo) + flink * k_mean + 52.0 * halo + 212.0
            + 128.0)


def host_cpus():
    """CPUs this process may use (the job's share on a shared GPU host: the
    OpenMP thread count the host sets, else the affinity mask) and the
    host's topology from lscpu."""
    avail = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(int(omp), avail) if omp.isdigit() and int(omp) > 0 else avail
    topo = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)",
                     "Model name"):
                topo[k] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    desc = (f"lscpu: {topo.get('Socket(s)', '?')} sockets x {topo.get('Core(s) per socket', '?')}"
            f" cores x {topo.get('Thread(s) per core', '?')} threads "
            f"({topo.get('CPU(s)', '?')} CPUs, {topo.get('Model name', '?')}); "
            f"{avail} CPUs in this process's affinity mask, OMP_NUM_THREADS={omp or 'unset'}")
    return threads, desc


def cpu_baseline(ns, ic, stencil, steps, threads, host):
    """The reference's own RK3 loop (oracle/_ref, flang/OpenMP, unmodified
    sources) on the full workload's node set, its neighbour lists and weights
    loaded from this library's (the reference's serial weight setup would
    take hours at this size).  The reference's LC right-hand side has no
    Kolmogorov line (rhs_lc.F90:263), so it runs the unforced flow; the
    forcing is one multiply-add per node.  Returns the baseline record and
    the reference's fields after 1 + steps steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refdrv
    if not refdrv.available(6):
        return None, None
    phys = {k: v for k, v in PHYS.items() if not k.startswith("kol")}
    out, times = refdrv.run_reference(ns, 6, ic, nsteps=steps + 1, phys=phys, stencil=stencil,
                                      fields_only=True, threads=threads, timeout=900)
    ts = sorted(times["steps"][1:])
    med = ts[len(ts) // 2]
    rec = dict(value=ns.npfb / med, unit="node-steps/s", cores=threads, kind="reference",
               sample=f"the full workload's {ns.npfb} nodes, same order/physics without the "
                      f"forcing line, {steps} timed steps after 1 warm-up, median step "
                      f"{med * 1e3:.1f} ms, OMP_NUM_THREADS={threads} OMP_PROC_BIND=close "
                      f"OMP_PLACES=cores",
               host=host)
    return rec, out


FP64_PEAK_TFLOPS = 78.6


def weight_flops(order, k_mean):
    """Algorithmic FP64 flops of one moment system (n unknowns, K links),
    counted for the textbook method whatever the device does: assembly
    2 K n^2 (MFMA), LU 2/3 n^3, three solves of 4 right-hand sides 24 n^2,
    two residuals 16 n^2, weights 8 K n.  (The device's Gauss-Jordan
    elimination does n^3 + 8 n^2 in place of the LU and the first solve.)"""
    n = (order * order + 3 * order) // 2
    return 2.0 * k_mean * n * n + 2.0 / 3.0 * n ** 3 + 40.0 * n * n + 8.0 * k_mean * n


def fp64_figure(flops, ms):
    tf = flops / (ms / 1e3) / 1e12
    return dict(bound="fp64", achieved=round(tf, 3), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                frac=round(tf / FP64_PEAK_TFLOPS, 4), flops=round(flops))


def c3_weights(smin, smax):
    """Weight setup on the cylinder-in-channel node set: spacing smin at the
    cylinder growing linearly to smax over 1.5, +/-0.3 s jitter, LABFM order 8
    inside and 4 within 3 s of any wall.  Times the one-block-per-system launch
    and the persistent work-queue launch (HIP events around the solves)."""
    import numpy as np
    import sunset_amd
    from sunset_amd.nodes import cylinder_channel
    ns, om = cylinder_channel(smin=smin, smax=smax)
    s = sunset_amd.Solver(morder=8, form="lc")
    s.set_nodes(ns)
    lists = s.find_neighbours()     # host copies for the oracle check below
    t_n = s.timing("neighbours")
    out = dict(nodes=ns.npfb, order8=int((om == 8).sum()), order4=int((om == 4).sum()),
               spacing=[smin, smax], neighbour_ms=round(t_n, 2))
    for name, kw in (("plain", dict(persistent=False)), ("persistent", dict(persistent=True)),
                     ("single_launch", dict(mixed=1)), ("autotuned", {})):
        s.calc_weights(order_map=om, **kw)
        t = s.timing("weights")
        out[f"{name}_ms"] = round(t, 2)
        out["systems"] = int(s.timing("systems"))
    mode = int(s.timing("weights_mixed"))
    out["autotuned_mode"] = ("per-order launches" if mode == 0 else
                             f"one persistent launch over all orders, {16 * mode}-neighbour tiles")
    out["autotune_mode_sample_ms"] = [round(s.timing(f"weights_mixed_tune_{k}"), 3)
                                      for k in range(3)]
    out["autotune_ms"] = round(s.timing("weights_tune"), 2)
    def variant_name(v):
        v = int(v)
        return (("persistent grid" if v & 1 else "one block per system")
                + (", 32-neighbour tiles" if v & 2 else ", 16-neighbour tiles"))
    out["autotuned_variants"] = {f"order{m}": variant_name(s.timing(f"weights_variant_m{m}"))
                                 for m in (4, 8)}
    out["autotune_sample_ms"] = {f"order{m}": [round(s.timing(f"weights_tune_m{m}_{v}"), 3)
                                               for v in range(4)] for m in (8,)}
    s.calc_sums()
    s.filter_coefficients()
    st = s.stencil_stats()
    out["error_vs_oracle"] = weights_error_vs_oracle(s, ns, om, lists)
    s.close()
    out.update(K_mean=round(st["K_mean"], 2), K_max=int(st["K_max"]),
               weight_solves_per_s=round(ns.npfb / (out["autotuned_ms"] / 1e3), 1))
    fl = (out["order8"] * weight_flops(8, st["K_mean"])
          + out["order4"] * weight_flops(4, st["K_mean"]))
    out["fp64"] = fp64_figure(fl, out["autotuned_ms"])
    return out


def weights_error_vs_oracle(s, ns, order_map, lists, samples=1000):
    """Largest per-node relative weight error (max_k |w - w_oracle| / max_k
    |w_oracle| over the derivative, Laplacian and filter weights) of a strided
    sample of nodes against the oracle's solve of those nodes on the same
    neighbour lists (one-sided Jacobi SVD; the reference's own Golub-Reinsch
    SVD is farther from the exact solution at order 8)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_lib as OL
    cnt, cs, link = lists
    nplink = link.shape[1]
    W = s.weights(nplink)
    o = OL.OracleSim(ns, 8, nplink=nplink)
    o.ghosts()
    o.load_lists(cnt, cs, link)
    del link
    n = ns.npfb
    sample = np.arange(0, n, max(1, n // samples), dtype=np.int32)[:samples]
    o.weights_subset(sample, order_map)
    OW = o.weight_arrays()
    worst = {}
    for key in ("w_grad", "w_lap", "w_hyp"):
        a = W[key][sample].reshape(sample.size, -1)
        b = OW[key][sample].reshape(sample.size, -1)
        den = np.abs(b).max(axis=1)
        worst[key] = float((np.abs(a - b).max(axis=1) / den).max())
    worst["vol"] = float(np.abs(W["vol"][sample] / OW["vol"][sample] - 1.0).max())
    return dict(max_rel=max(worst.values()), per_array=worst, sample_nodes=int(sample.size),
                how="strided node sample vs the oracle's solve on the device's lists")


def cpu_weights_baseline(n_sample):
    """The reference's calc_labf_weights (serial loop, labf.F90:82; derivative
    and filter systems, Golub-Reinsch SVDs) on an n_sample^2 order-8 lattice,
    h = 2.7 s; returns node weight sets per second."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refdrv
    from sunset_amd.nodes import jittered_lattice, taylor_green_ic
    if not refdrv.available(8):
        return None
    ns = jittered_lattice(n_sample, seed=2, h_over_s=2.7)
    _, times = refdrv.run_reference(ns, 8, taylor_green_ic(ns), nsteps=0, fields_only=True,
                                    threads=1, timeout=900)
    t = times["calc_labf_weights"]
    return dict(value=round(ns.npfb / t, 1), unit="node weight sets/s", cores=1,
                kind="reference",
                sample=f"{n_sample}x{n_sample} lattice ({ns.npfb} nodes), order 8, h=2.7s, "
                       f"calc_labf_weights {t:.2f} s (serial loop)")


def cpu_weights_omp(n_sample, threads):
    """The OpenMP-parallel CPU restatement of the weight solve (the oracle's
    calc_labf_weights with a per-node order, one node per thread, one-sided
    Jacobi SVDs) on the same order-8 lattice; node weight sets per second.
    The serial reference loop above is the x-baseline; this one shows what
    the host's cores give in parallel."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_lib as OL
    from sunset_amd.nodes import jittered_lattice
    ns = jittered_lattice(n_sample, seed=2, h_over_s=2.7)
    o = OL.OracleSim(ns, 8)
    o.ghosts()
    o.neighbours()
    om = np.full(ns.npfb, 8, dtype=np.int32)
    t0 = time.perf_counter()
    o.weights_ordermap(om, nthreads=threads)
    t = time.perf_counter() - t0
    return dict(value=round(ns.npfb / t, 1), unit="node weight sets/s", cores=threads,
                kind="port",
                sample=f"{n_sample}x{n_sample} lattice ({ns.npfb} nodes), order 8, h=2.7s, "
                       f"{t:.2f} s with {threads} OpenMP threads")


def boundary_leg(n=1024, steps=50, warmup=5):
    """A wall-bounded channel of the headline's size and physics: periodic in
    x, no-slip walls (characteristic wall conditions) at y = 0 and 1, FD rows
    along the normals, order 6, log-conformation sPTT, driven by a body force;
    the stage kernels take their boundary branch.  Node-steps/s over `steps`
    eager steps (HIP events) and the stage sweep's time per node."""
    import numpy as np
    import sunset_amd
    from sunset_amd.nodes import wall_channel
    ns = wall_channel(n, n, order=6, seed=2)
    s = sunset_amd.Solver(morder=6, form="lc").setup(ns)
    phys = {k: v for k, v in PHYS.items() if not k.startswith("kol")}
    s.set_phys(drive=(1.0, 0.0, 0.0), **phys)
    m = ns.npfb
    ic = dict(ro=np.ones(m), u=4 * ns.y * (1 - ns.y),
              v=0.01 * np.sin(2 * np.pi * ns.x) * np.sin(np.pi * ns.y) ** 2,
              cxx=np.ones(m), cxy=np.zeros(m), cyy=np.ones(m))
    s.set_initial(ic)
    s.run(warmup)
    kern, total_ms = s.run_profiled(steps)
    F = s.fields()
    s.close()
    if not all(np.isfinite(F[k]).all() for k in ("ro", "u", "v", "pxx", "pxy", "pyy")):
        raise SystemExit("bench.py: non-finite state in the wall-channel leg")
    t_stage = float(np.mean([kern[k] for k in ("stage1", "stage2", "stage3", "stage4")]))
    return dict(workload=f"{n}x{n} wall channel (walls at y=0,1, periodic in x), order 6, "
                         "log-conformation sPTT eps=0.25 Wi=2 beta=0.1, body force, PID dt",
                nodes=m, boundary_nodes=ns.nb, steps=steps,
                node_steps_per_s=round(m * steps / (total_ms / 1e3), 1),
                ms_per_step=round(total_ms / steps, 4), stage_ms=round(t_stage, 4))


def field_error(F, R, keys=("ro", "rou", "rov", "pxx", "pxy", "pyy", "cxx", "cxy", "cyy")):
    """max over the fields of max_i |F - R| / max(max_i |R|, 1)."""
    import numpy as np
    return max(float(np.abs(F[k] - R[k]).max() / max(np.abs(R[k]).max(), 1.0)) for k in keys)


def ensemble_leg(members=32, steps=100, seed=7):
    """32 simulations of different sizes in one context: N_m = round(exp(U(ln 16k,
    ln 256k))) nodes (square jittered lattices), LABFM order 4 or 6, Wi =
    exp(U(ln 0.1, ln 10)), beta = U(0.1, 0.9); log-conformation Oldroyd-B,
    Re = 1, Ma = 0.1, Kolmogorov forcing sin(4 pi y), PID dt.  Times `steps`
    steps of the whole ensemble (one hipGraph per step) against the same
    members run one after another (each its own graph)."""
    import numpy as np
    import sunset_amd
    from sunset_amd.nodes import ensemble_specs, jittered_lattice, taylor_green_ic
    specs = ensemble_specs(members, seed)

    def build():
        out = []
        for k, (side, order, Wi, beta) in enumerate(specs):
            print(f"[bench] ensemble member {k}: {side}^2 nodes, order {order}",
                  file=sys.stderr, flush=True)
            ns = jittered_lattice(side, seed=100 + k, h_over_s={4: 1.8, 6: 2.4}[order])
            s = sunset_amd.Solver(morder=order, form="lc").setup(ns)
            s.set_phys(Re=1.0, Wi=Wi, beta=beta, eps=0.0, Ma=0.1, kol_amp=1.0,
                       kol_k=4 * math.pi)
            s.set_initial(taylor_green_ic(ns))
            out.append(s)
        return out

    mem = build()
    nodes = sum(s.npfb for s in mem)
    print(f"[bench] ensemble: {members} members set up, {nodes} nodes; members one by one",
          file=sys.stderr, flush=True)
    t_seq = 0.0
    for s in mem:
        s.run(2, graph=True)
        s.run(steps, graph=True)
        t_seq += s.timing("step")
    print(f"[bench] ensemble: sequential {t_seq:.1f} ms; one ensemble context",
          file=sys.stderr, flush=True)
    E = sunset_amd.Ensemble(mem)   # copies the members' current states
    E.run(2, graph=True)
    E.run(steps, graph=True)
    t_ens = E.step_ms()
    mode = int(E.timing("ensemble_groups"))
    tune = [round(E.timing(f"ensemble_groups_tune_{k}"), 3) for k in (0, 1)]
    print(f"[bench] ensemble: {t_ens:.1f} ms for {steps} steps (launch mode {mode}, "
          f"tuning samples {tune} ms)", file=sys.stderr, flush=True)
    ok = all(np.isfinite(E.fields(k)["ro"]).all() for k in range(len(mem)))
    # the smallest member against its own oracle run over all its steps (the
    # one-by-one run above, then the ensemble's)
    km = min(range(len(mem)), key=lambda k: mem[k].npfb)
    side, order, Wi, beta = specs[km]
    F = E.fields(km)
    nsteps = int(F["itime"])
    E.close()
    for s in mem:
        s.close()
    if not ok:
        raise SystemExit("bench.py: non-finite ensemble state")
    print(f"[bench] ensemble: member {km} ({side}^2, order {order}) vs the oracle, "
          f"{nsteps} steps", file=sys.stderr, flush=True)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL
    ns = jittered_lattice(side, seed=100 + km, h_over_s={4: 1.8, 6: 2.4}[order])
    o = OL.OracleSim(ns, order).setup()
    o.make_phys(form="lc", Re=1.0, Wi=Wi, beta=beta, eps=0.0, Ma=0.1, kol_amp=1.0,
                kol_k=4 * math.pi)
    o.set_initial(taylor_green_ic(ns))
    o.run(nsteps)
    R = o.fields()
    err = field_error(F, R)
    return dict(members=members, nodes=nodes, steps=steps,
                sizes=[side * side for side, *_ in specs],
                orders=[o for _, o, _, _ in specs],
                ensemble_ms=round(t_ens, 2), sequential_ms=round(t_seq, 2),
                launch_mode=("one launch per order group" if mode == 1
                             else "one launch over all members"),
                launch_tune_ms=tune,
                node_steps_per_s=round(nodes * steps / (t_ens / 1e3), 1),
                sequential_node_steps_per_s=round(nodes * steps / (t_seq / 1e3), 1),
                speedup=round(t_seq / t_ens, 2),
                error_vs_oracle=dict(max_rel=err, member=km, nodes=ns.npfb, order=order,
                                     steps=nsteps,
                                     how="smallest member's fields vs its oracle run, "
                                         "max over fields of max|f - f_oracle| / max(|f|, 1)"))


def adjoint_leg(n=128, steps=400, every=20):
    """Checkpointed reverse mode on a 128^2 lattice, order 6, direct-form
    Oldroyd-B (Wi = 1, beta = 0.5, Re = 1, Ma = 0.1), fluid at rest plus a
    random-phase velocity perturbation of amplitude 0.01 (seed 5), Kolmogorov
    forcing sin(4 pi y), fixed dt = 2e-5; loss (1/T) sum_t mean(tr c - 2) over
    400 steps, a checkpoint every 20 steps.  Ratio = (forward rollout +
    checkpoint recompute + adjoint sweep) / forward rollout, device times;
    vjp_over_forward leaves the forward rollout out of the numerator."""
    import sunset_amd
    from sunset_amd.nodes import jittered_lattice, perturbed_rest_ic
    ns = jittered_lattice(n, seed=5, h_over_s=2.4)
    s = sunset_amd.Solver(morder=6, form="di").setup(ns)
    s.set_phys(Re=1.0, Wi=1.0, beta=0.5, eps=0.0, Ma=0.1, kol_amp=1.0, kol_k=4 * math.pi)
    ic = perturbed_rest_ic(ns, amp=0.01, seed=5)
    U0 = dict(ro=ic["ro"], rou=ic["ro"] * ic["u"], rov=ic["ro"] * ic["v"], cxx=ic["cxx"],
              cxy=ic["cxy"], cyy=ic["cyy"])
    s.set_state(U0, dt=2e-5)
    s.adj_setup(every)
    s.rollout_forward(steps)           # warm-up (allocations, first launches)
    s.rollout_vjp(1.0)
    s.set_state(U0, dt=2e-5)
    L = s.rollout_forward(steps)
    g = s.rollout_vjp(1.0)
    pg = s.adj_param_grad()
    fwd, bwd = s.adj_timing()
    s.close()
    import numpy as np
    gn = math.sqrt(sum(float(np.dot(v, v)) for v in g.values()))
    # initial velocity at fixed density: rou = ro*u, so dL/du = ro dL/drou
    gu = math.sqrt(float(np.dot(ic["ro"] * g["rou"], ic["ro"] * g["rou"])
                         + np.dot(ic["ro"] * g["rov"], ic["ro"] * g["rov"])))
    # the gradient along a random initial-velocity direction against the
    # oracle's tangent-linear rollout (forward-mode derivative of its FP64
    # step) over the same 400 steps
    print("[bench] adjoint: oracle tangent-linear rollout", file=sys.stderr, flush=True)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL
    n = ns.npfb
    rng = np.random.default_rng(21)
    du, dv = rng.standard_normal(n) * 0.01, rng.standard_normal(n) * 0.01
    dU = dict(ro=np.zeros(n), rou=ic["ro"] * du, rov=ic["ro"] * dv, cxx=np.zeros(n),
              cxy=np.zeros(n), cyy=np.zeros(n))
    dL = sum(float(np.dot(g[k], dU[k])) for k in dU)
    o = OL.OracleSim(ns, 6).setup()
    o.make_phys(form="di", dt_mode="fixed", Re=1.0, Wi=1.0, beta=0.5, eps=0.0, Ma=0.1,
                kol_amp=1.0, kol_k=4 * math.pi)
    o.set_initial(ic, dt=2e-5)
    Lo, dLo, _ = o.rollout_tlm(dU, steps)
    return dict(nodes=ns.npfb, steps=steps, checkpoint_every=every, loss=L, grad_norm=gn,
                grad_velocity_norm=gu, dL_dWi=pg["Wi"], dL_dbeta=pg["beta"],
                forward_ms=round(fwd, 2), vjp_ms=round(bwd, 2),
                ratio=round((fwd + bwd) / fwd, 3), vjp_over_forward=round(bwd / fwd, 3),
                error_vs_oracle=dict(
                    directional_derivative_rel=abs(dL - dLo) / abs(dLo),
                    loss_rel=abs(L - Lo) / abs(Lo), device=dL, oracle_tlm=dLo,
                    how="<grad, dU> along a random initial-velocity direction vs the "
                        "oracle's tangent-linear rollout, same 400 steps"))


def adapt_leg(n=512, check_n=96, cpu_n=48, no_cpu=False):
    """Stencil adaptation (adapt_stencils, labf.F90:1079-1721) of an n^2
    order-6 lattice on the device, list rebuild included; checked against the
    oracle in the device's arithmetic (correctly rounded Wendland power) on a
    check_n^2 lattice, and timed beside the reference's own serial loop on a
    cpu_n^2 lattice."""
    import numpy as np
    import sunset_amd
    from sunset_amd.nodes import jittered_lattice, taylor_green_ic
    ns = jittered_lattice(n, seed=3, h_over_s=2.4)
    s = sunset_amd.Solver(morder=6, form="lc")
    s.set_nodes(ns)
    s.find_neighbours(copy=False)
    t0 = time.perf_counter()
    h, hs, red = s.adapt_stencils()
    ms = (time.perf_counter() - t0) * 1e3
    s.close()
    out = {"nodes": ns.npfb, "order": 6, "ms": round(ms, 1),
           "nodes_per_s": round(ns.npfb / (ms / 1e3), 1),
           "mean_reductions": round(float(red[0].mean()), 3),
           "h_reduced_fraction": round(float((red[0] > 0).mean()), 4)}
    if no_cpu:
        return out
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL
    import refdrv
    nc = jittered_lattice(check_n, seed=77, h_over_s=2.4)
    s = sunset_amd.Solver(morder=6, form="lc")
    s.set_nodes(nc)
    s.find_neighbours(copy=False)
    dh, dhs, dred = s.adapt_stencils()
    s.close()
    with OL.wab_pow4(1):
        o = OL.OracleSim(nc, 6)
        o.ghosts()
        o.neighbours()
        ored = o.adapt_stencils()
        oh, ohs = o.h_hs()
    out["error_vs_oracle"] = {
        "nodes": nc.npfb, "equal_fraction": float(((dh == oh) & (dhs == ohs)).mean()),
        "max_reduction_difference": int(np.abs(dred - ored).max()),
        "how": "h and h_small of every node vs the oracle's adapt_stencils in the same "
               "arithmetic (correctly rounded (1-q/2)^4), 96^2 lattice"}
    if refdrv.available(6):
        ncp = jittered_lattice(cpu_n, seed=3, h_over_s=2.4)
        _, t = refdrv.run_reference(ncp, 6, taylor_green_ic(ncp), nsteps=0, adapt=True,
                                    threads=1, fields_only=True, timeout=600)
        sec = t.get("adapt_stencils")
        if sec:
            out["cpu_baseline"] = {
                "value": round(ncp.npfb / sec, 1), "unit": "nodes/s", "cores": 1,
                "kind": "reference",
                "sample": f"{cpu_n}^2 lattice ({ncp.npfb} nodes), order 6, adapt_stencils "
                          f"{sec:.2f} s (serial in the reference, labf.F90:1145-1148)"}
            out["x_vs_cpu"] = round(out["nodes_per_s"] / out["cpu_baseline"]["value"], 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1024, help="lattice side (n^2 nodes)")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU reference baseline and the oracle checks")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--graph", action="store_true", help="time hipGraph replay instead")
    ap.add_argument("--c3", choices=("full", "spec", "none"), default="full",
                    help="cylinder-channel weight setup: full = spacing 0.0013..0.0052 "
                         "(~1.2M nodes), spec = 0.004..0.016 (~126k nodes)")
    ap.add_argument("--cpu-weights-n", type=int, default=64)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-boundary", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-adapt", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is design-only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import sunset_amd
    from sunset_amd.nodes import jittered_lattice, taylor_green_ic

    ns = jittered_lattice(args.n, seed=2, h_over_s=2.4)
    print(f"[bench] {ns.npfb} nodes: setup", file=sys.stderr, flush=True)
    s = sunset_amd.Solver(morder=6, form="lc")
    setup = {}   # wall time of each setup call (host work included)

    def timed(name, f):
        t = time.perf_counter()
        r = f()
        setup[name] = round((time.perf_counter() - t) * 1e3, 1)
        return r

    timed("set_nodes", lambda: s.set_nodes(ns))
    cnt, cs, link = s.find_neighbours()    # host copies: CPU baseline and oracle check
    t_neigh = s.timing("neighbours")
    timed("calc_weights", s.calc_weights)
    t_w = s.timing("weights")
    n_sys = s.timing("systems")
    timed("calc_sums", s.calc_sums)
    timed("filter_coefficients", s.filter_coefficients)
    st = s.stencil_stats()   # halo statistics exist once the time loop has run
    s.set_phys(**PHYS)
    ic0 = taylor_green_ic(ns)
    timed("set_initial", lambda: s.set_initial(ic0))
    print(f"[bench] neighbours {t_neigh:.1f} ms, weights {t_w:.1f} ms "
          f"({n_sys:.0f} systems), K={st['K_mean']:.2f}; warmup {args.warmup}",
          file=sys.stderr, flush=True)
    s.run(args.warmup, graph=args.graph)
    s.fields()
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    if args.graph:
        t0 = time.perf_counter()
        s.run(args.steps, graph=True)
        total_ms = (time.perf_counter() - t0) * 1e3
        kern = None
    else:
        kern, total_ms = s.run_profiled(args.steps)
    F = s.fields()
    st = s.stencil_stats()
    if not all(np.isfinite(F[k]).all() for k in ("ro", "u", "v", "pxx", "pxy", "pyy")):
        raise SystemExit("bench.py: non-finite state after the timed steps")
    n = ns.npfb
    ms_step = total_ms / args.steps
    value = n * args.steps / (total_ms / 1e3)

    roof = None
    if kern is not None:
        t_stage = np.mean([kern[k] for k in ("stage1", "stage2", "stage3", "stage4")])
        b_stage = stage_bytes_per_node(st["Ks_mean"], st["halo_per_node"]) * n
        ach = b_stage / (t_stage / 1e3) / 1e9
        roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=measured_stage_traffic(),
                    kernel="k_stage_h (RK stage sweep, mean of stages 1-4)",
                    bytes_per_launch=int(b_stage), ms_per_launch=round(t_stage, 4))
    step_gbs = (step_bytes_per_node(st["Ks_mean"], st["K_mean"], st["halo_per_node"]) * n
                / (ms_step / 1e3) / 1e9)

    cpu = None
    c2_err = None
    threads, host = host_cpus()
    if not args.no_cpu:
        ic = taylor_green_ic(ns)
        stencil = dict(nplink=link.shape[1], ij_count=cnt, ij_count_small=cs, ij_link=link,
                       **s.weights(link.shape[1]))
        # one RK3 step from the initial state on the device and in the oracle
        # (run on the device's lists and weights), Kolmogorov forcing included
        print("[bench] error vs oracle: one step of the workload", file=sys.stderr, flush=True)
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as OL
        OL.lib().or_set_num_threads(threads)
        o = OL.OracleSim(ns, 6, nplink=link.shape[1])
        o.ghosts()
        o.load_lists(cnt, cs, link)
        o.load_weights(stencil)
        o.make_phys(form="lc", **PHYS)
        o.set_initial(ic)
        o.run(1)
        s.set_initial(ic)
        s.run(1)
        c2_err = dict(max_rel_one_step=field_error(s.fields(), o.fields()))
        del o
        print(f"[bench] cpu baseline: the reference on all {n} nodes, {threads} threads",
              file=sys.stderr, flush=True)
        try:
            cpu, ref = cpu_baseline(ns, ic, stencil, args.cpu_steps, threads, host)
        except (OSError, RuntimeError, subprocess.SubprocessError) as e:
            # the device figures stand without the CPU reference; say why it is missing
            cpu, ref = dict(value=None, error=str(e)[-400:], host=host), None
        del stencil
        if ref is not None:
            # the device on the reference's (unforced) flow, same steps
            s.set_phys(**{k: v for k, v in PHYS.items() if not k.startswith("kol")})
            s.set_initial(ic)
            s.run(args.cpu_steps + 1)
            c2_err["max_rel_vs_reference"] = field_error(s.fields(), ref)
            c2_err["reference_steps"] = args.cpu_steps + 1
        c2_err["how"] = ("max over fields of max|f - f_ref| / max(|f|, 1): one step vs the "
                         "oracle (forced flow); the reference's timed run (unforced) vs the "
                         "device's run of the same steps")
    del link

    c3 = None
    if args.c3 != "none":
        sp = (0.0013, 0.0052) if args.c3 == "full" else (0.004, 0.016)
        print(f"[bench] cylinder-channel weights, spacing {sp}", file=sys.stderr, flush=True)
        c3 = c3_weights(*sp)
        if not args.no_cpu:
            print(f"[bench] cpu weights baseline: reference on {args.cpu_weights_n}^2 order 8",
                  file=sys.stderr, flush=True)
            c3["cpu_baseline"] = cpu_weights_baseline(args.cpu_weights_n)
            print(f"[bench] cpu weights baseline: OpenMP port on {2 * args.cpu_weights_n}^2, "
                  f"{threads} threads", file=sys.stderr, flush=True)
            c3["cpu_baseline_omp"] = cpu_weights_omp(2 * args.cpu_weights_n, threads)
            c3["x_baseline"] = "cpu_baseline (serial reference loop)"
            if c3["cpu_baseline"]:
                c3["x_vs_cpu_serial"] = round(c3["weight_solves_per_s"]
                                              / c3["cpu_baseline"]["value"], 1)
            c3["x_vs_cpu_omp"] = round(c3["weight_solves_per_s"]
                                       / c3["cpu_baseline_omp"]["value"], 1)

    bnd = None
    if not args.no_boundary:
        print("[bench] wall channel (boundary nodes)", file=sys.stderr, flush=True)
        bnd = boundary_leg()
        bnd["vs_periodic"] = round(bnd["node_steps_per_s"] / value, 3)

    ens = None
    if not args.no_ensemble:
        print("[bench] ensemble: 32 members", file=sys.stderr, flush=True)
        ens = ensemble_leg()
    adj = None
    if not args.no_adjoint:
        print("[bench] adjoint: 128^2 DI rollout of 400 steps", file=sys.stderr, flush=True)
        adj = adjoint_leg()

    adapt = None
    if not args.no_adapt:
        print("[bench] stencil adaptation: 512^2 order 6", file=sys.stderr, flush=True)
        adapt = adapt_leg(no_cpu=args.no_cpu)

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "node-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded jittered lattice, Taylor-Green initial velocity)",
        "config": {"workload": "1024^2 jittered periodic lattice, LABFM order 6, h=2.4s, "
                               "log-conformation sPTT eps=0.25 Wi=2 beta=0.1 Re=1 Ma=0.1, "
                               "Kolmogorov f_x=sin(4 pi y), PID dt",
                   "nodes": n, "K_mean": round(st["K_mean"], 2), "order": 6,
                   "halo_records_per_node": round(st["halo_per_node"], 3),
                   "timing": "hipGraph replay" if args.graph else "eager, HIP events"},
        "node_stage_updates_per_s": round(4 * value, 1),
        "cpu_node_stage_updates_per_s": (round(4 * cpu["value"], 1)
                                         if cpu and cpu.get("value") else None),
        "roofline": roof,
        "cpu_baseline": cpu,
        "error_vs_oracle": c2_err,
        "gpu_vs_cpu": round(value / cpu["value"], 1) if cpu and cpu.get("value") else None,
        "step_hbm_gbs": round(step_gbs, 1),
        "weight_solves_per_s": round(n / (t_w / 1e3), 1),
        "weights_fp64": fp64_figure(n_sys * weight_flops(6, st["Ks_mean"]), t_w),
        "cylinder_channel_weights": c3,
        "weight_setup_ms": round(t_w, 2), "neighbour_ms": round(t_neigh, 2),
        "setup_wall_ms": setup,
        "kernel_ms": {k: round(v, 4) for k, v in kern.items()} if kern else None,
        "wall_channel": bnd,
        "ensemble": ens,
        "adjoint": adj,
        "stencil_adaptation": adapt,
        "adjoint_forward_ratio": adj["ratio"] if adj else None,
        "adjoint_vjp_over_forward": adj["vjp_over_forward"] if adj else None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

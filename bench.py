"""Throughput benchmark of the explicit MPM step on one MI355X.

Workload: 2D Quadrilateral4Node 1024x1024 cells on [0,1]^2, 4 particles per
cell at jittered (+-0.25 h, seed 0) Gauss points: 4,194,304 particles,
LinearElastic E=1e6, nu=0.3, rho=1000, gravity (0,-9.81), bottom nodes fixed,
dt=1e-5, USF, FP32.  W warm-up steps, then K timed steps bracketed by
synchronisation (the throughput); the step inputs are resident in HBM.  The
next K steps are timed again with HIP events around every kernel class for
the per-kernel durations of the roofline.

Prints one JSON line with particle-steps/s, the roofline of the dominant
kernel (HIP-event durations measured inside the second timed region) and the CPU
baseline (the oracle, a NumPy restatement of the reference step, timed on a
bounded strip of the same workload on this host).
"""
import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per particle of each timed kernel class (FP32 2D
# LinearElastic): the persistent particle fields the kernel must read and
# write; grid arrays, sort keys and recomputable temporaries excluded.
KERNEL_BYTES = {
    # sort by cell: sort keys and permutation only (no persistent particle field)
    "sort": 0,
    # P2G of mass and momentum (cell + node pass), nodal velocity, constraints:
    # read x(2), v(2), m
    "tile_kin": 4 * (2 + 2 + 1),
    # USF: strain/volume/stress update from the kinematic nodal
    # velocities: read x(2), eps(3), sigma(4), V, rho, eps_v; write eps(3),
    # sigma(4), V, rho, eps_v
    "stress": 4 * (2 + 3 + 4 + 3) + 4 * (3 + 4 + 3),
    # P2G of body and internal forces (cell + node pass), nodal
    # forces, nodal update, constraints: read x(2), m, V, sigma(xx, yy, xy)
    "tile_force": 4 * (2 + 1 + 1 + 3),
    # G2P: read x(2), v(2); write x(2), v(2)
    "tile_g2p": 4 * 4 + 4 * 4,
    # the one-launch step (default): every persistent field of the step once,
    # read x(2), v(2), m, V, rho, sigma(4), eps(3), eps_v; write all but m
    "step": 4 * (2 + 2 + 1 + 1 + 1 + 4 + 3 + 1) + 4 * (2 + 2 + 1 + 1 + 4 + 3 + 1),
}
# --usf-fused: the force cell pass also does the stress update in place: read
# x(2), m, eps(3), sigma(4), V, rho, eps_v; write eps(3), sigma(4), V, rho, eps_v
TILE_FORCE_FUSED_BYTES = 4 * (2 + 1 + 3 + 4 + 3) + 4 * (3 + 4 + 3)
STEP_BYTES = 116  # per particle-step, whole step (FP32 2D LinearElastic)
# adjoint step: read primal x(2), v(2), m, V, sigma(xx, yy, xy); read and write
# the cotangents of x(2), v(2), V, sigma(xx, yy, xy); write dL/dlogE_p
ADJOINT_BYTES = 4 * 9 + 4 * 8 * 2 + 4
# HBM bytes per particle per launch of each kernel class at 1024x1024 (step: the
# default one-launch path; the others: --phase-launches), from rocprofv3 --pmc
# passes of this benchmark: reads = TCC_EA0_RDREQ
# x 128 B, writes = WRITE_SIZE (calibrated on a 512 MiB device copy; see
# profiles/step_q4_1024_usf_hbm_counters.md)
TRAFFIC_BYTES = {"step": 135.8, "sort": 13.5, "tile_kin": 57.7, "stress": 91.0,
                 "tile_force": 86.4, "tile_g2p": 38.1}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _strip_sample(rows):
    """The throughput workload's bottom `rows` cell rows as a 1024 x `rows`
    mesh: the same jittered particles (clamped below the strip's top node
    row), bottom nodes fixed in x and y, side rollers."""
    import numpy as np
    from diffmpm_amd.workloads import gauss_points
    n = 1024
    h = 1.0 / n
    pts = gauss_points(n, rows, h, 0.25, 0)
    top = float(np.nextafter(np.float32(rows) * np.float32(h), np.float32(0.0)))
    np.minimum(pts[:, 1], top, out=pts[:, 1])
    b, j = np.arange(n + 1), np.arange(rows + 1)
    cons = [(b, 0, 0.0), (b, 1, 0.0), (np.concatenate([j * (n + 1), j * (n + 1) + n]), 0, 0.0)]
    return n, h, pts, cons


def cpu_baseline(rows=64, np_steps=12, torch_steps=80):
    """The CPU restatements of the reference step on a 1024 x `rows`-cell
    strip of the throughput workload (same particles, constraints, material,
    gravity, dt, USF): the NumPy oracle on one core, and the torch restatement
    of the same algorithm on every host thread (about 10 s each)."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import mpm as om
    from oracle.torch_step import TorchQ4
    try:
        from threadpoolctl import threadpool_limits
        lim = threadpool_limits(1)
    except Exception:
        lim = None
    n, h, pts, cons = _strip_sample(rows)
    props = {"density": 1000.0, "youngs_modulus": 1e6, "poisson_ratio": 0.3}
    g = np.array([0.0, -9.81])
    el = om.Elements("quad4", (n, rows), (h, h),
                     [(ids, om.Constraint(d, v)) for ids, d, v in cons], [], np.float32)
    ps = om.Particles(pts, om.LinearElastic(props), dtype=np.float32)
    sim = om.MPMExplicit(om.Mesh(el, [ps]), 1e-5, "usf", False, 1)
    sim.step(0, g)  # warm-up
    t0 = time.perf_counter()
    for i in range(1, np_steps + 1):
        sim.step(i, g)
    np_s = time.perf_counter() - t0
    if lim is not None and hasattr(lim, "unregister"):
        lim.unregister()
    assert np.isfinite(ps.loc).all() and np.isfinite(ps.stress).all()
    threads = len(os.sched_getaffinity(0))
    torch.set_num_threads(threads)
    de = om.LinearElastic(props).de64
    t = TorchQ4(n, rows, h, h, torch.from_numpy(pts.astype(np.float32)), de, 1000.0, cons, 1e-5)
    t.step(0, g)
    t0 = time.perf_counter()
    for i in range(1, torch_steps + 1):
        t.step(i, g)
    th_s = time.perf_counter() - t0
    assert torch.isfinite(t.loc).all() and torch.isfinite(t.stress).all()
    sample = (f"1024x{rows}-cell strip of the throughput workload ({len(pts)} particles: "
              f"its bottom {rows} cell rows, the same jittered positions, bottom fixed, side "
              f"rollers, E=1e6, nu=0.3, rho=1000, gravity, dt=1e-5, USF, FP32)")
    return {"value": len(pts) * torch_steps / th_s, "unit": "particle-steps/s",
            "cores": threads, "kind": "port", "cpu_model": _cpu_model(),
            "sample": f"torch CPU restatement of the reference step (oracle/torch_step.py), "
                      f"{threads} threads, {torch_steps} steps after 1 warm-up "
                      f"({th_s:.1f} s), on a {sample}",
            "config": {"workload": "Q4 strip", "nx": n, "ny": rows, "particles": len(pts),
                       "scheme": "usf", "constraints": "bottom x+y fixed, side rollers"},
            "single_core": {"value": len(pts) * np_steps / np_s, "unit": "particle-steps/s",
                            "cores": 1, "kind": "port",
                            "sample": f"NumPy oracle (np.add.at scatters in particle order, "
                                      f"1 thread), {np_steps} steps after 1 warm-up "
                                      f"({np_s:.1f} s), same strip"}}


def ensemble_bench(K, K_adj=None, K_seq_adj=50, checkpoint_every=100):
    """64 uneven USL members (32..256 cells per side, 30-100 % of the rows
    filled, particles at the Gauss points).  Forward: K steps with one batch
    plan (the autotuner's choice of step kernels, the K steps captured into
    one HIP graph and replayed; the eager run of the same plan beside it),
    against the same K steps of every member with its own plan, one member
    after the other.  Forward+adjoint w.r.t. each member's scalar E (one log
    E per member, mapped to its particles): K_adj (default K) batched steps with the
    adjoint checkpointed every checkpoint_every steps, against K_seq_adj
    steps of every member's own differentiable rollout, one after the other
    (per-step rates; the sequential rollouts are bound by their per-step host
    work, so fewer steps time them)."""
    import math
    import time
    import torch
    from diffmpm_amd import workloads
    from diffmpm_amd._plan import plan_for
    from diffmpm_amd.autodiff import Differentiable, rollout
    from diffmpm_amd.ensemble import Ensemble
    from diffmpm_amd.graphs import StepGraph
    sims, g, specs = workloads.ensemble(64, seed=0)
    ntot = sum(len(s.mesh.particles[0]) for s in sims)
    plans = [plan_for(s.mesh.elements, s.mesh.particles[0]) for s in sims]
    for p, s in zip(plans, sims):  # warm-up (workspaces, first full sort)
        p.run("usl", False, g, s.dt, 0, 2, keep_temporaries=False)
    ens = Ensemble(sims)
    tuned = ens.autotune(g)
    ens.run(2, g, start_step=2)
    graph = StepGraph(ens, g, K, start_step=4)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record(stream)
    graph.replay()  # steps 4 .. 4+K-1
    ev[1].record(stream)
    ens.run(K, g, start_step=4 + K)
    ev[2].record(stream)
    for p, s in zip(plans, sims):
        p.run("usl", False, g, s.dt, 4, K, keep_temporaries=False)
    ev[3].record(stream)
    torch.cuda.synchronize()
    tg, te, ts = (ev[i].elapsed_time(ev[i + 1]) for i in range(3))
    assert torch.isfinite(ens.pbuf[0:2, :ens.n]).all()
    nodes = graph.kernel_nodes()
    del graph
    out = {"workload": "64 uneven Q4 members, n = 32..256 cells per side (log-uniform), "
                       "30-100 % rows filled at the Gauss points, E 1e5..1e7, nu 0.2..0.4, "
                       "USL, dt=1e-5, seed 0",
           "particles": ntot, "steps": K,
           "batched": {"value": ntot * K / (tg / 1e3), "unit": "particle-steps/s",
                       "ms_per_step": tg / K, "run": "HIP graph of the K steps, one replay",
                       "graph_kernel_nodes_per_step": nodes / K},
           "batched_eager": {"value": ntot * K / (te / 1e3), "unit": "particle-steps/s",
                             "ms_per_step": te / K},
           "sequential": {"value": ntot * K / (ts / 1e3), "unit": "particle-steps/s",
                          "ms_per_step": ts / K},
           "speedup": ts / tg,
           "autotune_ms_per_step": {str(k): v for k, v in tuned.items()},
           "batch_step": "one launch" if ens.step_kernel else
                         "per phase, %d cells per workgroup" % ens.cells_per_block}
    K_adj = K if K_adj is None else K_adj
    K_seq_adj = min(K_seq_adj, K_adj)
    if not K_adj:
        return out
    # forward+adjoint: fresh members, per-member scalar log E
    del ens
    torch.cuda.empty_cache()
    sims, g, specs = workloads.ensemble(64, seed=0)
    ens = Ensemble(sims, per_particle_E=True)
    ens.autotune(g)
    d = Differentiable.of_ensemble(ens)
    idx = ens.sim_index()
    le0 = torch.tensor([math.log(m["E"]) for m in specs], device="cuda")

    def batched(k):
        le = le0.clone().requires_grad_(True)
        ck = checkpoint_every if k > checkpoint_every else 0
        o = rollout(d, k, g, state=d.get_state(), log_e=le[idx], checkpoint_every=ck,
                    graphs=bool(ck))
        o[0:2].sum().backward()
        return le.grad

    batched(4)  # warm-up
    batched(2 * checkpoint_every)  # captures the segments' graphs
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    grad = batched(K_adj)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    assert torch.isfinite(grad).all()
    del d, ens
    torch.cuda.empty_cache()
    sims, g, specs = workloads.ensemble(64, seed=0)
    ds = [Differentiable(s) for s in sims]
    for dd in ds:  # warm-up
        le = torch.tensor(math.log(dd.ps.youngs_modulus[0].item()), device="cuda",
                          requires_grad=True)
        rollout(dd, 2, g, log_e=le)[0:2].sum().backward()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for dd in ds:
        le = torch.tensor(math.log(dd.ps.youngs_modulus[0].item()), device="cuda",
                          requires_grad=True)
        rollout(dd, K_seq_adj, g, log_e=le)[0:2].sum().backward()
    torch.cuda.synchronize()
    tq = time.perf_counter() - t0
    out["fwd_adjoint"] = {
        "batched": {"value": ntot * K_adj / tb, "unit": "particle-steps/s", "steps": K_adj,
                    "checkpoint_every": checkpoint_every, "ms_per_step": tb * 1e3 / K_adj,
                    "run": "per segment: one HIP graph replay forward, one backward "
                           "(recomputation with saved adjoint inputs, then the adjoint steps)"},
        "sequential": {"value": ntot * K_seq_adj / tq, "unit": "particle-steps/s",
                       "steps": K_seq_adj, "ms_per_step": tq * 1e3 / K_seq_adj},
        "speedup": (ntot * K_adj / tb) / (ntot * K_seq_adj / tq),
        "parameter": "d loss / d log E of each member (one scalar per member), loss = sum of "
                     "final positions"}
    return out


def fwd_adjoint(sim, gravity, K):
    """Throughput of a differentiable rollout: K forward steps recorded for
    autograd (each keeps what its adjoint needs), then the backward pass (the
    adjoint kernels per step), per-particle log E as the parameter."""
    import math
    import torch
    from diffmpm_amd.autodiff import Differentiable, rollout
    from diffmpm_amd._plan import timing_enable, timing_read
    d = Differentiable(sim)
    N = d.n
    log_e = torch.full((N,), math.log(1e6), dtype=d.ps.dtype, device=d.ps.device,
                       requires_grad=True)
    state0 = d.get_state().requires_grad_(True)
    # warm-up of the same length (fills the allocator's cache for the saved
    # states and cotangents)
    out = rollout(d, K, gravity, state=state0, log_e=log_e)
    out[0:2].sum().backward()
    log_e.grad = None
    state0.grad = None
    torch.cuda.synchronize()
    aplan = d.fplan if d.save_forward else d.aplan  # the plan the adjoint runs on
    stream = torch.cuda.current_stream()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))

    def timed():
        log_e.grad = None
        state0.grad = None
        torch.cuda.synchronize()
        e0.record(stream)
        out = rollout(d, K, gravity, state=state0, log_e=log_e)
        e1.record(stream)
        out[0:2].sum().backward()
        e2.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1), e1.elapsed_time(e2)

    # throughput: no events between the kernels; then the adjoint kernels'
    # durations from a second rollout with HIP events around them
    fwd_ms, bwd_ms = timed()
    timing_enable(aplan, "adj")
    timed()
    adj_ms, adj_n = timing_read(aplan)
    timing_enable(aplan, None)
    assert torch.isfinite(log_e.grad).all() and torch.isfinite(state0.grad).all()
    ms = (fwd_ms + bwd_ms) / K
    adj_us = adj_ms / max(adj_n, 1) * 1e3
    return {"metric": "particle-steps/s fwd+adjoint at 4.2M ptcl FP32",
            "value": N / (ms / 1e3), "unit": "particle-steps/s", "steps": K,
            "ms_per_step": ms, "forward_ms_per_step": fwd_ms / K,
            "backward_ms_per_step": bwd_ms / K,
            "adjoint_kernels_us_per_step": adj_us,
            "adjoint_kernels_GBs": ADJOINT_BYTES * N / (adj_us / 1e6) / 1e9,
            "note": "forward steps (one-launch step between two blocks, state rows in the "
                    "autograd tensors) keep the node block, kinematic nodal velocity and sort "
                    "order their adjoint needs; backward = adjoint kernels; per-particle log "
                    "E; loss = sum of final positions"}


def inverse_bench(iters, K=2000, checkpoint_every=100):
    """The differentiable inverse problem at its configuration size
    (workloads.inverse_2d(128): 32,768 particles, USF, dt = 1e-5): the
    target is the final positions of a K-step rollout with the true E_p
    field; Adam (lr 1e-2 on log E) from the uniform E = 1e6 minimises the
    mean square position error (divided by the target's mean square
    displacement, so that Adam's epsilon stays below the gradient scale),
    each iteration a K-step rollout with the adjoint checkpointed every
    checkpoint_every steps.  Reports the time per iteration, the
    forward+adjoint step rate, and the loss and log E error trajectory."""
    import math
    import time
    import torch
    from diffmpm_amd import workloads
    from diffmpm_amd.autodiff import Differentiable, rollout
    sim, g, E_true = workloads.inverse_2d(128)
    d = Differentiable(sim)
    N = d.n
    state0 = d.get_state().clone()
    le_true = torch.log(E_true.to("cuda", torch.float32))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.no_grad():
        target = rollout(d, K, g, state=state0, log_e=le_true)[0:2].clone()
    torch.cuda.synchronize()
    fwd_s = time.perf_counter() - t0
    scale = ((target - state0[0:2]) ** 2).mean()
    le = torch.full((N,), math.log(1e6), device="cuda", requires_grad=True)
    opt = torch.optim.Adam([le], lr=1e-2)
    dev = (le_true - le_true.mean()).norm()
    losses, errs, secs = [], [], []
    for _ in range(iters):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        opt.zero_grad()
        out = rollout(d, K, g, state=state0, log_e=le, checkpoint_every=checkpoint_every,
                      graphs=True)
        loss = ((out[0:2] - target) ** 2).mean() / scale
        loss.backward()
        opt.step()
        torch.cuda.synchronize()
        secs.append(time.perf_counter() - t0)
        losses.append(loss.item())
        errs.append(float((le.detach() - le_true).norm() / dev))
    it_s = sorted(secs)[len(secs) // 2]
    return {"workload": "inverse problem: Q4 128x128, 32768 particles (lower half, Gauss "
                        "points), E_p = 1e6 exp(0.3 g), g Gaussian field (corr. 0.1, seed 0), "
                        "nu 0.3, rho 1000, gravity, bottom fixed, USF, dt 1e-5",
            "particles": N, "steps": K, "checkpoint_every": checkpoint_every,
            "iterations": iters, "s_per_iteration": it_s,
            "fwd_adjoint_particle_steps_per_s": N * K / it_s,
            "forward_only_s": fwd_s, "loss": losses, "log_e_rel_error": errs,
            "note": "one iteration = forward rollout (autograd records segment starts), "
                    "backward (each segment recomputed, then its adjoint steps; each "
                    "segment's steps one HIP graph replay), Adam step; "
                    "median time over the iterations"}


def copy_ceiling(nbytes=1 << 30, reps=10):
    """STREAM-copy ceiling of this GPU: a device-to-device copy of nbytes,
    timed with HIP events over reps copies after a warm-up; GB/s of read
    plus written bytes."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    b.copy_(a)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        b.copy_(a)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * nbytes / (ms / 1e3) / 1e9


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--kernel", default="auto",
                    help="kernel class of the roofline (auto: the one with the most time)")
    ap.add_argument("--cells-per-block", type=int, default=0,
                    help="cells per workgroup of the P2G cell pass (64/128/256; 0: default)")
    ap.add_argument("--usf-fused", action="store_true",
                    help="phase launches: update stresses inside the force P2G cell pass")
    ap.add_argument("--phase-launches", action="store_true",
                    help="one launch per phase instead of the one-launch step")
    ap.add_argument("--step-kernel", choices=["one", "split"], default="one",
                    help="the step's kernels: one launch per step, or the split step "
                         "(three launches over node and cell tiles)")
    ap.add_argument("--ensemble-steps", type=int, default=1000,
                    help="steps of the timed 64-member uneven ensemble (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=20,
                    help="steps of the timed forward+adjoint rollout (0: skip)")
    ap.add_argument("--inverse-iters", type=int, default=20,
                    help="Adam iterations of the 2000-step 128x128 inverse problem (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed region, copy 512 MiB with a plain device copy "
                         "(reference kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    import torch
    from diffmpm_amd import workloads
    from diffmpm_amd._plan import plan_for, timing_enable, timing_read

    sim, gravity = workloads.block_2d(args.n)
    ps = sim.mesh.particles[0]
    plan = plan_for(sim.mesh.elements, ps)
    N = len(ps)
    if args.cells_per_block:
        plan.set_block(args.cells_per_block)
    from diffmpm_amd import _lib
    if args.usf_fused:
        plan.set_option(_lib.DMPM_OPT_USF_SPLIT, 0)
    if args.phase_launches or args.usf_fused:
        plan.set_option(_lib.DMPM_OPT_STEP_KERNEL, 0)
    elif args.step_kernel == "split":
        plan.set_option(_lib.DMPM_OPT_STEP_KERNEL, 2)
    print(f"# workload built: {N} particles, {args.n}x{args.n} cells", file=sys.stderr, flush=True)
    plan.run("usf", False, gravity, sim.dt, 0, args.warmup, keep_temporaries=False)
    torch.cuda.synchronize()
    print("# warm-up done", file=sys.stderr, flush=True)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(step0):
        torch.cuda.synchronize()
        e0.record(stream)
        plan.run("usf", False, gravity, sim.dt, step0, args.steps, keep_temporaries=False)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    # throughput: K steps with no events between the kernels
    ms = timed(args.warmup)
    # per-kernel durations: the next K steps with HIP events recorded around
    # each kernel class on the launch stream (the events add gaps between
    # launches, so this region is not the one the throughput comes from)
    timing_enable(plan, "all")
    ms_instr = timed(args.warmup + args.steps)
    kbytes_tab = dict(KERNEL_BYTES)
    if args.usf_fused:
        kbytes_tab["tile_force"] = TILE_FORCE_FUSED_BYTES
    kernels = {}
    for kc in kbytes_tab:
        kms, kn = timing_read(plan, kc)
        if kn:
            kernels[kc] = {"avg_us": kms / kn * 1e3, "launches": kn,
                           "bytes_per_particle": kbytes_tab[kc],
                           "GBs": kbytes_tab[kc] * N / (kms / kn / 1e3) / 1e9}
    timing_enable(plan, None)
    assert torch.isfinite(ps.loc).all() and torch.isfinite(ps.stress).all()
    value = N * args.steps / (ms / 1e3)
    if args.kernel == "auto":
        cand = {k: v for k, v in kernels.items() if v["bytes_per_particle"]}
        kname = max(cand, key=lambda k: cand[k]["avg_us"] * cand[k]["launches"])
    else:
        kname = args.kernel
    kd = kernels[kname]
    achieved = kd["GBs"]
    traffic = TRAFFIC_BYTES.get(kname)
    line = {
        "metric": "particle-steps/s fwd at 4.2M ptcl FP32",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"Q4 {args.n}x{args.n} cells, 4 ppc jittered, "
                               f"{N} particles, USF, LinearElastic, gravity, dt=1e-5",
                   "particles": N, "scheme": "usf"},
        "step_effective_GBs": STEP_BYTES * N * args.steps / (ms / 1e3) / 1e9,
        "ms_per_step_with_kernel_events": ms_instr / args.steps,
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (traffic * N) if (traffic and args.n == 1024
                                                   and not args.usf_fused
                                                   and args.step_kernel == "one") else None,
                     "avg_launch_us": kd["avg_us"],
                     "bytes_per_particle": kd["bytes_per_particle"], "launches": kd["launches"]},
        "kernels": kernels,
    }
    ceil = copy_ceiling()
    line["roofline"]["copy_ceiling_GBs"] = ceil
    line["roofline"]["frac_of_copy_ceiling"] = achieved / ceil
    if args.adjoint_steps:
        line["fwd_adjoint"] = fwd_adjoint(sim, gravity, args.adjoint_steps)
    if args.ensemble_steps:
        del sim, ps, plan
        torch.cuda.empty_cache()
        line["ensemble"] = ensemble_bench(args.ensemble_steps)
    if args.inverse_iters:
        line["inverse"] = inverse_bench(args.inverse_iters)
    if args.copy_calibration:
        a = torch.empty(128 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
    if not args.no_cpu_baseline:
        print("# cpu baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

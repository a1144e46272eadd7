"""Benchmark of the nbfmm force-evaluation + Verlet step on one MI355X.

A step is one nbfmm::Solver::solve plus the Verlet update (Stars::update
without prune) over the whole system.  Default workload: the two-disk
collision, N = 2^20 particles, 512x512 leaves (num_level 8) over [-4,4]^2,
eps 1e-3, dt 1e-4, seeded synthetic initial conditions.

Prints ONE JSON line with the throughput in particle-steps/s, the P2P
kernel's roofline figures (HIP events around the P2P launches of an
instrumented pass over the same steps) and the CPU oracle timed on a bounded
sample of the same workload as a baseline.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X published FP32 vector peak (packed FP32)
FP32_PEAK_NONPACKED_TFLOPS = 78.6  # the same without packed FP32, which this build does not use
HBM_PEAK_GBS = 8000.0
P2P_FLOPS_PER_PAIR = 14    # sub 2, r^2 3, sqrt 1, r^3 2, +eps 1, div 1, scale 2, accumulate 2
# Pair-rate bound of the P2P arithmetic on MI355X: a throughput probe running
# the inner loop's instruction mix per pair (8 FMA-class VALU + v_sqrt_f32 +
# v_rcp_f32, independent chains, every CU busy) sustains 3.19e12 lane-steps/s,
# one pair per lane-step (8 FMA alone: 4.45e12; transcendentals cost ~1.9 FMA
# slots each and mostly add to the rest).
P2P_PAIR_BOUND_PER_S = 3.19e12
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_two_disk_N1048576.md")


def measured_traffic(kernel="k_p2p<kP2PStepSort>"):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (None if absent)."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                return float(line.split(":")[1])
    except OSError:
        return None
    return None


# Algorithmic (compulsory) HBM bytes of the other stages of a step, per
# particle, per leaf and per tree cell (all levels):
#   sort (warm, from its second pass): read the binned particles of the
#     previous order (key 4 + particle 16 + rank 4) and write the sorted key
#     and particle (4 + 16): 44 B per particle; head and the occupied-leaf
#     list: 8 B per leaf;
#   upward (P2M + M2M): read x, y, m (12 B per particle); write every cell's
#     multipole (16 B) and read it once more from its parent (16 B, all but
#     the top level);
#   M2L: read each cell's multipole (16 B) and write its local (8 B).
SORT_B_PER_PARTICLE, SORT_B_PER_LEAF = 44, 8
UP_B_PER_PARTICLE, UP_B_PER_CELL = 12, 32
M2L_B_PER_CELL = 24
M2L_FLOPS_PER_PAIR = 14


def tree_cells(num_level):
    B = 1 << (num_level + 1)
    return sum((B >> l) ** 2 for l in range(max(num_level, 1)))


def m2l_pairs(num_level):
    """M2L cell pairs of one step (source/nbfmm/core/solver/m2l.cu:39-51):
    per level and axis, the window [p-2, p+4) of the target's parent p
    (clipped to the grid) minus its own 3-cell neighbourhood."""
    B = 1 << (num_level + 1)
    tot = 0
    for l in range(num_level):
        D = B >> l
        W = sum(min((t & ~1) + 4, D) - max((t & ~1) - 2, 0) for t in range(D))
        Nn = sum(min(t + 2, D) - max(t - 1, 0) for t in range(D))
        tot += W * W - Nn * Nn
    return tot


def log(msg):
    print(msg, file=sys.stderr, flush=True)


import numpy as np  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="two_disk_collision")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU baseline sample length")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-systems leg")
    ap.add_argument("--batch-systems", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint legs")
    ap.add_argument("--no-small", action="store_true", help="skip the 16384-particle disk leg")
    ap.add_argument("--no-tune", action="store_true", help="keep the default P2P configuration")
    ap.add_argument("--no-sensitivity", action="store_true",
                    help="skip the step at a larger G and the cold-sort step")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one 256 MiB device-to-device copy (scale reference for HBM counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the multi-GPU sharding is a design "
                                   "(spatial decomposition with a halo exchange), not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    from nbfmm_amd import Solver
    from nbfmm_amd.workloads import by_name

    wl = by_name(args.workload)
    n = wl.n
    log(f"[bench] workload {wl.name} N={n} num_level={wl.num_level} steps={args.steps} warmup={args.warmup}")
    dev = torch.device("cuda:0")
    s = Solver(wl.num_level, n, wl.limits, wl.eps)
    x = torch.as_tensor(wl.position, device=dev).contiguous()
    p = torch.as_tensor(wl.position_pre, device=dev).contiguous()
    w = torch.as_tensor(wl.weight, device=dev).contiguous()
    a = torch.zeros_like(x)
    x0, p0 = x.clone(), p.clone()
    stream = torch.cuda.Stream()

    # P2P configuration: timed on this input (all configurations give the
    # same bits), then the pair count at the initial state (the P2P work figure)
    tuned = None
    if not args.no_tune:
        from nbfmm_amd.autotune import tune, variants
        with torch.cuda.stream(stream):
            v, bpc, tms = tune(s, n, x, w, cache_path=None, repeats=3)
        stage, unroll, waves = variants()[v]
        tuned = {"variant": v, "stage": stage, "unroll": unroll, "waves_per_block": waves, "blocks_per_cu": bpc,
                 "p2p_ms": tms}
        log(f"[bench] tuned P2P: {tuned}")
    with torch.cuda.stream(stream):
        s.solve(n, x, w, a)
    stream.synchronize()
    pairs0, occ0 = s.pair_count(n)

    # warmup + timed region: hipGraph replay of the step
    with torch.cuda.stream(stream):
        s.run(n, wl.tick, args.warmup, x, p, w, a, use_graph=True)
    stream.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        s.run(n, wl.tick, args.steps, x, p, w, a, use_graph=True)
    stream.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    dt_s = t1 - t0
    value = n * args.steps / dt_s
    pairs1, occ1 = s.pair_count(n)
    flags = s.flags()
    log(f"[bench] graph: {dt_s * 1e3 / args.steps:.4f} ms/step, {value:.4e} particle-steps/s, flags={flags}")

    # instrumented pass over the same number of steps from the same state,
    # eager, with the launches the graph replays (each step's P2P is
    # k_p2p<kP2PStepSort>, which also makes the next sort's first pass, but
    # the last one's): HIP events on the launch stream around each stage
    x.copy_(x0)
    p.copy_(p0)
    torch.cuda.synchronize()
    s.profile(True)
    s.stage_times(reset=True)
    te0 = time.perf_counter()
    with torch.cuda.stream(stream):
        s.run(n, wl.tick, args.steps, x, p, w, a, use_graph=False)
    stream.synchronize()
    te1 = time.perf_counter()
    stage_ms, launches = s.stage_times(reset=True)
    s.profile(False)
    avg_ms = [t / max(launches, 1) for t in stage_ms]
    p2p_avg_ms = avg_ms[3]
    pairs = 0.5 * (pairs0 + pairs1)
    achieved = P2P_FLOPS_PER_PAIR * pairs / (p2p_avg_ms * 1e-3) / 1e12
    log(f"[bench] eager: {(te1 - te0) * 1e3 / args.steps:.4f} ms/step; stages (us) "
        f"{[round(t * 1e3, 1) for t in avg_ms]}; P2P pairs/launch {pairs:.4e}; {achieved:.2f} TFLOP/s")
    B = 1 << (wl.num_level + 1)
    cells = tree_cells(wl.num_level)
    npairs_m2l = m2l_pairs(wl.num_level)

    def hbm(nbytes, ms):
        gbs = nbytes / (ms * 1e-3) / 1e9
        return {"us": ms * 1e3, "algorithmic_bytes": nbytes, "GBps": gbs, "frac_hbm": gbs / HBM_PEAK_GBS}

    stages = {
        "sort": hbm(SORT_B_PER_PARTICLE * n + SORT_B_PER_LEAF * B * B, avg_ms[0]),
        "upward": hbm(UP_B_PER_PARTICLE * n + UP_B_PER_CELL * cells, avg_ms[1]),
        "m2l": hbm(M2L_B_PER_CELL * cells, avg_ms[2]),
        "p2p": {"us": p2p_avg_ms * 1e3, "flop": P2P_FLOPS_PER_PAIR * pairs, "TFLOPs": achieved,
                "frac_fp32": achieved / FP32_PEAK_TFLOPS},
        "note": "eager launches of the replayed step, HIP events between stages (launch gaps included); "
                "sort = warm passes 2-5 (the first is inside the previous step's P2P)",
    }
    stages["m2l"]["flop"] = M2L_FLOPS_PER_PAIR * npairs_m2l
    stages["m2l"]["TFLOPs"] = M2L_FLOPS_PER_PAIR * npairs_m2l / (avg_ms[2] * 1e-3) / 1e12
    sens = None if args.no_sensitivity else sort_sensitivity(args, wl)

    accuracy = None
    cpu = None
    if not args.no_accuracy or not args.no_cpu:
        from oracle import pyoracle as orc
        lim = np.array(wl.limits, np.float32)
    if not args.no_accuracy:
        # force parity (GPU vs oracle) and the FMM error against the fp64 direct sum
        ta = time.perf_counter()
        with torch.cuda.stream(stream):
            s.solve(n, x0, w, a)
        stream.synchronize()
        g = a.cpu().numpy().astype(np.float64)
        ref, _ = orc.solve(wl.num_level, wl.position, wl.weight, lim, wl.eps)
        rng = np.random.default_rng(0)
        tg = np.sort(rng.choice(n, size=min(n, 1024), replace=False))
        direct = orc.direct_f64(wl.position, wl.weight, wl.eps, tg)
        accuracy = {
            "gpu_vs_oracle_rel_l2": float(np.linalg.norm(g - ref) / np.linalg.norm(ref)),
            "gpu_vs_direct_fp64_rel_l2": float(np.linalg.norm(g[tg] - direct) / np.linalg.norm(direct)),
            "oracle_vs_direct_fp64_rel_l2": float(np.linalg.norm(ref[tg] - direct) / np.linalg.norm(direct)),
            "direct_targets": int(tg.size),
        }
        log(f"[bench] accuracy {accuracy} ({time.perf_counter() - ta:.1f} s)")

    if not args.no_cpu:
        # CPU oracle (C++/OpenMP restatement of the reference algorithm) on a
        # bounded sample: whole steps of the same workload, all host threads.
        threads = orc.max_threads()
        tc0 = time.perf_counter()
        orc.run(wl.num_level, lim, wl.eps, wl.tick, 1, wl.position, wl.position_pre, wl.weight)
        one = time.perf_counter() - tc0
        ksteps = int(max(1, min(400, args.cpu_seconds // max(one, 1e-3))))
        tc0 = time.perf_counter()
        orc.run(wl.num_level, lim, wl.eps, wl.tick, ksteps, wl.position, wl.position_pre, wl.weight)
        tcpu = time.perf_counter() - tc0
        cpu = {"value": n * ksteps / tcpu, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": f"{ksteps} full steps of {wl.name} N={n} (after 1 untimed step), OpenMP, "
                         f"{tcpu:.1f} s",
               "cpu_model": orc.cpu_model(), "isa": orc.isa(), "nproc": os.cpu_count()}
        log(f"[bench] cpu baseline {cpu}")

    if args.calib_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    small = None if args.no_small else small_leg()
    batched = None if args.no_batch else batch_leg(args)
    adjoint = None if args.no_adjoint else adjoint_leg()
    batched_adjoint = None if (args.no_adjoint or args.no_batch) else batch_adjoint_leg()

    out = {
        "metric": "particle-steps/s at N=2^20; batched+adjoint vs plain, force rel err <=1e-4",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt_s * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl.name, "n": n, "num_level": wl.num_level, "leaf_grid": 1 << (wl.num_level + 1),
                   "limits": list(wl.limits), "eps": wl.eps, "dt": wl.dt, "G": wl.G, "mode": "plain, hipGraph",
                   "parallelism": "single GPU", "p2p_config": tuned},
        "roofline": {"kernel": "k_p2p<kP2PStepSort>", "bound": "valu", "achieved": achieved,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "peak_nonpacked": FP32_PEAK_NONPACKED_TFLOPS,
                     "frac_nonpacked": achieved / FP32_PEAK_NONPACKED_TFLOPS,
                     "interactions_per_s": pairs / (p2p_avg_ms * 1e-3),
                     "traffic": (measured_traffic() if wl.name == "two_disk_collision" and n == 1 << 20
                                 else None),
                     "avg_launch_us": p2p_avg_ms * 1e3, "launches": launches,
                     "pairs_per_launch": pairs, "flops_per_pair": P2P_FLOPS_PER_PAIR,
                     "valu_issue_frac": pairs / (p2p_avg_ms * 1e-3 * P2P_PAIR_BOUND_PER_S),
                     "occupied_leaves": int(occ0)},
        "stages": stages,
        "sort_sensitivity": sens,
        "cpu_baseline": cpu,
        "accuracy": accuracy,
        "small_disk": small,
        "batched": batched,
        "adjoint": adjoint,
        "batched_adjoint": batched_adjoint,
    }
    s.close()
    print(json.dumps(out), flush=True)


def sort_sensitivity(args, wl0):
    """What a step costs when particles change leaf more often: the same
    workload at G = 1e4 / N (velocities 100x, about 100x the leaf changes per
    step, warm-started sort) and at the default G with the cold sort every
    step (warm start off), both graph-replayed like the headline."""
    import torch

    from nbfmm_amd import Solver
    from nbfmm_amd.workloads import two_disk_collision

    out = {}
    steps = max(10, args.steps // 2)
    for key, G, warm in (("fast_G_1e4_over_N", 1e4 / wl0.n, True), ("cold_sort", None, False)):
        wl = two_disk_collision(n=wl0.n, num_level=wl0.num_level, G=G)
        s = Solver(wl.num_level, wl.n, wl.limits, wl.eps)
        s.warm_start(warm)
        x = torch.as_tensor(wl.position, device="cuda").contiguous()
        p = torch.as_tensor(wl.position_pre, device="cuda").contiguous()
        w = torch.as_tensor(wl.weight, device="cuda").contiguous()
        a = torch.zeros_like(x)
        st = torch.cuda.Stream()
        with torch.cuda.stream(st):
            s.run(wl.n, wl.tick, 3, x, p, w, a, use_graph=True)
        st.synchronize()
        t0 = time.perf_counter()
        with torch.cuda.stream(st):
            s.run(wl.n, wl.tick, steps, x, p, w, a, use_graph=True)
        st.synchronize()
        dt = time.perf_counter() - t0
        out[key] = {"G": wl.G, "warm_start": warm, "steps": steps, "ms_per_step": dt * 1e3 / steps,
                    "value": wl.n * steps / dt, "flags": int(s.flags())}
        s.close()
    log(f"[bench] sort sensitivity {out}")
    return out


def batch_leg(args):
    """Many independent systems (N log-uniform in [1024, 131072], uneven
    depths): one hipGraph replay per step with every stage one launch over
    all systems, against the same systems stepped one Solver at a time."""
    import torch

    from nbfmm_amd import Batch, Solver
    from nbfmm_amd.workloads import many_disks

    bw = many_disks(num_systems=args.batch_systems, steps=args.batch_steps)
    steps = bw.steps
    total = bw.n
    log(f"[bench] batch: {len(bw.systems)} systems, {total} particles, {steps} steps")
    x, p, w = bw.concat()
    dev = torch.device("cuda:0")
    X = torch.as_tensor(x, device=dev)
    P = torch.as_tensor(p, device=dev)
    W = torch.as_tensor(w, device=dev)
    A = torch.zeros_like(X)
    b = Batch([s.n for s in bw.systems], [s.num_level for s in bw.systems], [s.limits for s in bw.systems],
              [s.eps for s in bw.systems], ticks=[s.tick for s in bw.systems])
    stream = torch.cuda.Stream()
    X0, P0 = X.clone(), P.clone()

    def batched_run():
        # the same `steps` steps from the same start each time (the graphs,
        # captured on the first call, read X and P in place)
        X.copy_(X0)
        P.copy_(P0)
        with torch.cuda.stream(stream):
            b.run(2, X, P, W, A, use_graph=True)
        stream.synchronize()
        t0 = time.perf_counter()
        with torch.cuda.stream(stream):
            b.run(steps, X, P, W, A, use_graph=True)
        stream.synchronize()
        return time.perf_counter() - t0

    # the GPU's clocks ramp up under load: about a second of these runs
    # first, then the median of 3 timed repeats
    t_end = time.perf_counter() + 1.0
    while time.perf_counter() < t_end:
        batched_run()
    reps = [batched_run() for _ in range(3)]
    tb = sorted(reps)[1]
    log(f"[bench] batch repeats (ms per step): {[round(t * 1e3 / steps, 4) for t in reps]}")
    flags = int(b.flags().any())
    b.close()
    # per-system launches: one Solver per system, stepped one after the other
    solvers = [Solver(s.num_level, s.n, s.limits, s.eps) for s in bw.systems]
    off = np.concatenate([[0], np.cumsum([s.n for s in bw.systems])])
    views = [(X[off[k]:off[k + 1]], P[off[k]:off[k + 1]], W[off[k]:off[k + 1]], A[off[k]:off[k + 1]])
             for k in range(len(bw.systems))]
    with torch.cuda.stream(stream):
        for k, s in enumerate(bw.systems):
            solvers[k].step(s.n, s.tick, *views[k])
    stream.synchronize()
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        for _ in range(steps):
            for k, s in enumerate(bw.systems):
                solvers[k].step(s.n, s.tick, *views[k])
    stream.synchronize()
    ts = time.perf_counter() - t0
    for sv in solvers:
        sv.close()
    out = {"workload": bw.name, "systems": len(bw.systems), "particles": total, "steps": steps,
           "batched_value": total * steps / tb, "per_system_value": total * steps / ts,
           "speedup": ts / tb, "unit": "particle-steps/s", "batched_ms_per_step": tb * 1e3 / steps,
           "per_system_ms_per_step": ts * 1e3 / steps, "flags": flags}
    log(f"[bench] batch {out}")
    return out


def small_leg(reps=20):
    """Plain FMM on the 16384-particle uniform disk (32x32 leaves over
    [-1.1,1.1]^2, eps 1e-3, dt 1e-4): its 10 Verlet steps replayed as a
    hipGraph, each of `reps` timed runs starting again from the initial state
    (two 128 KiB copies on the same stream, inside the timed region); force
    parity against the oracle and the FMM error against the full fp64 direct
    sum at the initial state.  At this size a step is a few microseconds of
    work per launch, so the figure mostly measures launch latency."""
    import torch

    from nbfmm_amd import Solver
    from nbfmm_amd.workloads import disk
    from oracle import pyoracle as orc

    wl = disk()
    n = wl.n
    dev = torch.device("cuda:0")
    s = Solver(wl.num_level, n, wl.limits, wl.eps)
    x = torch.as_tensor(wl.position, device=dev).contiguous()
    p = torch.as_tensor(wl.position_pre, device=dev).contiguous()
    w = torch.as_tensor(wl.weight, device=dev).contiguous()
    a = torch.zeros_like(x)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        s.solve(n, x, w, a)
    stream.synchronize()
    g = a.cpu().numpy().astype(np.float64)
    x0, p0 = x.clone(), p.clone()
    with torch.cuda.stream(stream):
        s.run(n, wl.tick, wl.steps, x, p, w, a, use_graph=True)  # warm-up (captures the graph)
    stream.synchronize()
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        for _ in range(reps):
            x.copy_(x0)
            p.copy_(p0)
            s.run(n, wl.tick, wl.steps, x, p, w, a, use_graph=True)
    stream.synchronize()
    dt = time.perf_counter() - t0
    flags = s.flags()
    s.close()
    lim = np.array(wl.limits, np.float32)
    ref, _ = orc.solve(wl.num_level, wl.position, wl.weight, lim, wl.eps)
    direct = orc.direct_f64(wl.position, wl.weight, wl.eps)
    out = {"workload": wl.name, "n": n, "num_level": wl.num_level, "steps_per_run": wl.steps, "runs": reps,
           "value": n * wl.steps * reps / dt, "unit": "particle-steps/s", "ms_per_step": dt * 1e3 / (wl.steps * reps),
           "gpu_vs_oracle_rel_l2": float(np.linalg.norm(g - ref) / np.linalg.norm(ref)),
           "gpu_vs_direct_fp64_rel_l2": float(np.linalg.norm(g - direct) / np.linalg.norm(direct)),
           "oracle_vs_direct_fp64_rel_l2": float(np.linalg.norm(ref - direct) / np.linalg.norm(direct)),
           "flags": int(flags)}
    log(f"[bench] small disk {out}")
    return out


def adjoint_leg():
    """Differentiable trajectory: 65536-particle Gaussian disk, 128x128
    leaves, 50 Verlet steps checkpointed every 10; loss = mean squared
    distance of the final positions to targets from a run with masses
    perturbed by 5% Gaussian noise; gradients wrt initial positions, x_pre
    and masses.  Reports forward and backward time."""
    import torch

    from nbfmm_amd import Solver
    from nbfmm_amd.autograd import fmm_trajectory
    from nbfmm_amd.workloads import gaussian_disk_system

    wl = gaussian_disk_system()
    n, steps, K = wl.n, wl.steps, 10
    dev = torch.device("cuda:0")
    s = Solver(wl.num_level, n, wl.limits, wl.eps)
    x0 = torch.as_tensor(wl.position, device=dev)
    p0 = torch.as_tensor(wl.position_pre, device=dev)
    m_true = torch.as_tensor(wl.weight, device=dev)
    rng = np.random.default_rng(31)
    m_pert = m_true * torch.as_tensor((1 + 0.05 * rng.normal(size=n)).astype(np.float32), device=dev)
    with torch.no_grad():
        target = fmm_trajectory(s, wl.tick, steps, K, [steps], x0, p0, m_pert)

    def fwd():
        xg = x0.clone().requires_grad_(True)
        pg = p0.clone().requires_grad_(True)
        mg = m_true.clone().requires_grad_(True)
        out = fmm_trajectory(s, wl.tick, steps, K, [steps], xg, pg, mg)
        loss = ((out - target) ** 2).sum(-1).mean()
        return loss, (xg, pg, mg)

    loss, leaves = fwd()  # warm-up
    loss.backward()
    torch.cuda.synchronize()
    reps = 3
    tf = tb = 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        loss, leaves = fwd()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        tf += t1 - t0
        tb += t2 - t1
    tf /= reps
    tb /= reps
    gm = leaves[2].grad
    s.close()
    sig = signal_over_rounding(float(loss.detach()), target)
    out = {"workload": wl.name, "n": n, "num_level": wl.num_level, "steps": steps, "checkpoint_every": K,
           "G": wl.G, "mass_signal": sig,
           "forward_ms": tf * 1e3, "backward_ms": tb * 1e3, "backward_over_forward": tb / tf,
           "forward_particle_steps_per_s": n * steps / tf,
           "forward_backward_particle_steps_per_s": n * steps / (tf + tb),
           "loss": float(loss.detach()), "grad_mass_finite": bool(torch.isfinite(gm).all())}
    log(f"[bench] adjoint {out}")
    return out


def signal_over_rounding(loss, target):
    """The fit's signal against fp32 rounding: the RMS position mismatch at
    the starting masses (sqrt of the mean squared distance) over the RMS
    fp32 spacing (one ulp) of the observed positions."""
    t = target.detach().float().cpu().numpy()
    ulp = float(np.sqrt(np.mean(np.spacing(np.abs(t).astype(np.float32)) ** 2)))
    rms = float(np.sqrt(max(loss, 0.0)))
    return {"rms_mismatch": rms, "rms_fp32_ulp": ulp, "ratio": rms / ulp if ulp > 0 else None}


def batch_adjoint_leg(reps=3):
    """Batched mass inference: 64 two-disk collisions (N log-uniform in
    [4096, 262144]), 100 steps checkpointed every 10, loss = position
    mismatch at steps 25/50/75/100 against a run with the true masses,
    starting masses = true x U[0.8,1.2]; gradients wrt the masses of every
    system in one batched backward.  Forward and backward times are means
    over `reps` repetitions (after the untimed run that makes the target)."""
    import torch

    from nbfmm_amd import Batch
    from nbfmm_amd.autograd import fmm_batch_trajectory
    from nbfmm_amd.workloads import FIT_G_PER_N, many_collisions

    bw = many_collisions()
    steps, K, obs = bw.steps, 10, [25, 50, 75, 100]
    x, p, w = bw.concat()
    dev = torch.device("cuda:0")
    X, P, W = (torch.as_tensor(a, device=dev) for a in (x, p, w))
    b = Batch([s.n for s in bw.systems], [s.num_level for s in bw.systems], [s.limits for s in bw.systems],
              [s.eps for s in bw.systems], ticks=[s.tick for s in bw.systems])
    rng = np.random.default_rng(41)
    with torch.no_grad():
        target = fmm_batch_trajectory(b, steps, K, obs, X, P, W)
    torch.cuda.synchronize()
    m0 = W * torch.as_tensor(rng.uniform(0.8, 1.2, W.shape[0]).astype(np.float32), device=dev)
    tf = tb = 0.0
    for _ in range(reps):
        m = m0.clone().requires_grad_(True)
        t0 = time.perf_counter()
        out = fmm_batch_trajectory(b, steps, K, obs, X, P, m)
        loss = ((out - target) ** 2).sum(-1).mean()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        tf += t1 - t0
        tb += t2 - t1
    tf /= reps
    tb /= reps
    b.close()
    total = bw.n
    # the loss at the starting masses (the last repetition's; all start alike)
    sig = signal_over_rounding(float(loss.detach()), target)
    res = {"workload": bw.name, "systems": len(bw.systems), "particles": total, "steps": steps,
           "checkpoint_every": K, "repeats": reps, "forward_ms": tf * 1e3, "backward_ms": tb * 1e3,
           "backward_over_forward": tb / tf,
           "forward_backward_particle_steps_per_s": total * steps / (tf + tb), "loss": float(loss.detach()),
           "grad_mass_finite": bool(torch.isfinite(m.grad).all()), "G_per_N": FIT_G_PER_N,
           "mass_signal": sig}
    log(f"[bench] batched adjoint {res}")
    return res


if __name__ == "__main__":
    main()

"""Force-path benchmark on one MI355X.

Workload (default ``c2``): periodic unit box, 128^3 = 2,097,152 particles on a
grid displaced by the Zel'dovich field of P(k) ~ k^-2.5 (rms |psi| = 0.1 grid
spacing, Philox seed 7), Plummer softening 1/40 grid spacing, relative
opening criterion alpha = 0.005, 65^3 FP64 Ewald table.  A step is one
domain decomposition (keys, sort, reorder, top nodes) plus
Compute_Acceleration on a regular step (tree build, relative FP64 walk,
Ewald walk), all on device-resident data.

Output: one JSON line.  ``value`` = force evaluations per second of the
whole step (N_active / t_step).  ``roofline`` covers the dominant kernel,
timed with HIP events on the engine's stream, counting 20 flops per
monopole interaction (and 60 per Ewald interaction).  ``cpu_baseline``
times the CPU oracle (the C/OpenMP restatement of the reference path) on
the host cores: decomposition and build on all particles, the walks on a
bounded random sample of active particles, scaled to the same unit.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"f64": 78.6e12, "f32": 157.3e12, "hbm": 8.0e12}
FLOPS_PER_INTERACTION = 20.0
FLOPS_PER_EWALD = 60.0

WORKLOADS = {
    "c2": dict(ngrid=128, seed=7, alpha=0.005, desc="periodic Zel'dovich box "
               "128^3, FP64 walk + Ewald (drop-in path)"),
    "c2_small": dict(ngrid=64, seed=7, alpha=0.005, desc="periodic "
                     "Zel'dovich box 64^3, FP64 walk + Ewald"),
}


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def make_workload(name):
    from tandav_amd import ics
    w = WORKLOADS[name]
    pos, vel, m = ics.zeldovich_box(w["ngrid"], seed=w["seed"])
    soft = 1.0 / (40 * w["ngrid"])
    return pos, vel, m, soft, w


def run_gpu(pos, m, soft, w, steps, warmup):
    from tandav_amd import Engine
    e = Engine(periodic=True, boxsize=1.0, softening=soft, theta_bh=0.1,
               alpha_rel=w["alpha"], compat=True)
    e.set_particles(pos, m)
    e.setup_domain()
    e.domain_decomposition()
    log("prepare step (BH walk, then relative walk)")
    e.compute_acceleration(prepare_step=True)       # BH walk, then relative
    for i in range(warmup):
        e.domain_decomposition()
        e.compute_acceleration(prepare_step=False)
        log("warmup step %d: %s" % (i, e.stage_times()))
    stage = {k: 0.0 for k in e.stage_times()}
    t0 = time.perf_counter()
    for _ in range(steps):
        e.domain_decomposition()
        e.compute_acceleration(prepare_step=False)
        for k, v in e.stage_times().items():
            stage[k] += v
        log("timed step done")
    t1 = time.perf_counter()
    cost = e.get("cost")["cost"]
    return e, (t1 - t0) / steps, {k: v / steps for k, v in stage.items()}, cost


def cpu_baseline(pos, m, soft, w, sample, nthreads, gpu_acc, gpu_cost):
    """Oracle on the host cores: decomposition + build on every particle,
    then the steady-state walks (relative + Ewald) on a random sample of
    particles with the same previous-step |a| the GPU step used, so the
    sample does exactly the GPU's work (its Cost is checked against the
    GPU's) and scales to the whole step."""
    import oracle
    n = len(m)
    table = oracle.cached_ewald_table()
    s = oracle.Sim(pos, m, periodic=True, boxsize=1.0, softening=soft,
                   theta_bh=0.1, alpha_rel=w["alpha"], nthreads=nthreads,
                   ewald_table=table)
    s.setup_domain()
    t0 = time.perf_counter()
    s.domain_decomposition()
    t1 = time.perf_counter()
    log("cpu: decomposition %.2fs" % (t1 - t0))
    s.tree_build()
    t2 = time.perf_counter()
    log("cpu: tree build %.2fs" % (t2 - t1))
    g = np.random.Generator(np.random.Philox(99))
    pick = np.sort(g.choice(n, sample, replace=False))
    s.time_bin[:] = 1
    s.time_bin[pick] = 0
    s.set_active_bin(0)
    s.last_acc_mag[:] = gpu_acc          # same PH order as the GPU arrays
    t3 = time.perf_counter()
    s.tree_acceleration(use_bh=False)
    t4 = time.perf_counter()
    log("cpu: relative+Ewald walk on %d particles %.2fs" % (sample, t4 - t3))
    same_cost = bool(np.array_equal(s.cost[pick], gpu_cost[pick]))
    t_step = (t1 - t0) + (t2 - t1) + (t4 - t3) * n / sample
    return {"value": n / t_step, "unit": "force-evals/s", "cores": nthreads,
            "kind": "port",
            "sample": "oracle (C99+OpenMP restatement of the reference path): "
                      "decomposition %.2fs + tree build %.2fs on all %d "
                      "particles, steady-state relative+Ewald walk on %d "
                      "random particles %.2fs scaled by N/sample; sample "
                      "Cost identical to the GPU's: %s"
                      % (t1 - t0, t2 - t1, n, sample, t4 - t3, same_cost)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-sample", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus %d is not "
              "supported (the multi-GPU sharding is a design note only)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    log("generating %s" % args.workload)
    pos, vel, m, soft, w = make_workload(args.workload)
    n = len(m)
    log("N = %d" % n)
    e, t_step, stage, cost = run_gpu(pos, m, soft, w, args.steps, args.warmup)
    last_cost = cost.copy()
    ms_step = t_step * 1e3

    # dominant kernel and its roofline (HIP events on the engine's stream)
    walk_ms, ewald_ms = stage["walk"], stage["ewald"]
    tree_inter, ewald_inter = e.interaction_counts()
    if walk_ms >= ewald_ms:
        kern, t_k = "k_walk64", walk_ms
        flops = tree_inter * FLOPS_PER_INTERACTION
    else:
        kern, t_k = "k_walk64_ewald", ewald_ms
        flops = ewald_inter * FLOPS_PER_EWALD
    achieved = flops / (t_k * 1e-3)
    roof = {"bound": "fp64-valu", "kernel": kern,
            "achieved": achieved / 1e12, "peak": PEAK["f64"] / 1e12,
            "unit": "TFLOP/s", "frac": achieved / PEAK["f64"],
            "traffic": None,
            "flops_per_interaction": FLOPS_PER_INTERACTION
            if kern == "k_walk64" else FLOPS_PER_EWALD,
            "interactions_per_launch": tree_inter if kern == "k_walk64"
            else ewald_inter,
            "avg_launch_ms": t_k}

    cpu = None
    if not args.no_cpu_baseline:
        # the |a| the GPU's last step opened nodes with (Last_Acc_Mag,
        # accel.c:28-43) and the Cost that step produced
        g = e.get("last_acc_mag")
        cpu = cpu_baseline(pos, m, soft, w, min(args.cpu_sample, n),
                           min(os.cpu_count() or 1, 16),
                           g["last_acc_mag"], last_cost)

    walk_s = (walk_ms + ewald_ms) * 1e-3
    line = {
        "metric": "tree-force evals/s per step, FP32 vs FP64 walk, at p99 "
                  "|da|/|a| <= 2e-3",
        "value": n / t_step,
        "unit": "force-evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": w["desc"], "n_particles": n,
                   "softening": soft, "alpha_rel": w["alpha"],
                   "walk": "fp64", "parallelism": "single GPU"},
        "stages_ms": {k: round(v, 4) for k, v in stage.items()},
        "walk_force_evals_per_s": n / walk_s,
        "interactions_per_s": float(cost.sum()) / walk_s,
        "interactions_per_particle": float(cost.mean()),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

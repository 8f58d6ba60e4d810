"""Force-path benchmark on one MI355X: FP32 group walk vs FP64 walk.

Workload (default ``c3``): periodic unit box, 256^3 = 16,777,216 particles on
a grid displaced by the Zel'dovich field of P(k) ~ k^-2.5 (rms |psi| = 0.1
grid spacing, Philox seed 7), Plummer softening 1/40 grid spacing, 65^3
Ewald table (``c2``: the same at 128^3).  A step is one domain decomposition
(PH keys, radix sort, reorder, top nodes) plus Compute_Acceleration on a
regular step (octree build, tree walk with the Ewald correction), all on
device-resident data.  The step runs with the FP64 walk (the drop-in path:
the reference's walk and interaction sets, relative criterion 0.005) and
with the FP32 group walk (monopole + quadrupole nodes, second-order Ewald
correction at centres of mass, the next-order relative criterion
M s^3 <= r^5 |a| alpha / G) at alpha 0.01, then 0.005 and 0.0025.

Output: one JSON line.  ``value`` = force evaluations per second of the
whole step with the FP32 group walk at alpha 0.01 (N_active / t_step);
``fp64`` holds the same for the FP64 walk, ``fp32_alpha_scan`` the other
criteria.  Accuracy: |da|/|a| of every walk against FP64 direct summation
over all particles with the exact periodic force (the Ewald lattice sum
per pair, no table) on 512 sampled particles (the error oracle; the bar is
p99 <= 2e-3), and against Gravity_Forcetest's form (the same sum with the
reference's 65^3 table, trilinear) on 8192, whose own error against the
exact force is reported too.  ``roofline`` covers the dominant kernel,
timed with HIP events on the engine's stream, counting the flops each
interaction's formula takes (FLOPS below: 20 for a point mass, 50 for a
node with its quadrupole, 60 for an Ewald correction, 120 with the
quadrupole term), interactions counted on the device.  ``cpu_baseline`` times
the CPU oracle (the C/OpenMP restatement of the reference path) on the
host cores: decomposition and build on all particles, the FP64 walks on a
bounded sample of particles.
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
# The FP32 walk's interaction kinds and the flops of their formulas:
#   point mass   d (3), r^2 (5), r^-1 (1), m r^-3 (3), a += f d (6),
#                the softening test (1), carry share (1)              20
#   node + quad  d (3), r^2 (5), r^-1 (1), u = d r^-1 (3), T u (15),
#                u.T u (5), the scalars r^-2, r^-4, g (6),
#                a += g u - r^-4 T u (12)                              50
#   Ewald        node lookup (13), second-order expansion (42), m E (6) 60
#   Ewald + quad the same plus T:S/6 (6 scalings, 18 fma pairs, 6)    120
FLOPS = {"point": 20.0, "node_quad": 50.0, "ewald": 60.0, "ewald_quad": 120.0}
COUNTING = ("flops per interaction by kind: 20 point mass, 50 node with "
            "quadrupole, 60 Ewald correction, 120 Ewald correction with "
            "quadrupole; interactions counted on the device")


def walk32_flops(kinds):
    """Flops of one FP32 walk from tdv_interaction_kinds."""
    inter, ewald, nodes, ewq = kinds
    return (FLOPS["point"] * (inter - nodes) + FLOPS["node_quad"] * nodes
            + FLOPS["ewald"] * (ewald - ewq) + FLOPS["ewald_quad"] * ewq)

WORKLOADS = {
    "c2": dict(ngrid=128, seed=7, alpha=0.005, desc="periodic Zel'dovich box "
               "128^3 (2,097,152 particles), Ewald-corrected"),
    "c2_small": dict(ngrid=64, seed=7, alpha=0.005, desc="periodic "
                     "Zel'dovich box 64^3, Ewald-corrected"),
    "c3": dict(ngrid=256, seed=7, alpha=0.005, desc="periodic Zel'dovich box "
               "256^3 (16,777,216 particles), Ewald-corrected"),
}
# the FP32 group walk's settings (the FP64 walk keeps the reference's)
FP32 = dict(alpha=0.01, fp32_criterion=3, fp32_order=2, ewald_size=0.125,
            ewald_accuracy=1.0)


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def make_workload(name):
    from tandav_amd import ics
    w = WORKLOADS[name]
    pos, vel, m = ics.zeldovich_box(w["ngrid"], seed=w["seed"])
    soft = 1.0 / (40 * w["ngrid"])
    return pos, vel, m, soft, w


def run_gpu(pos, m, soft, w, steps, warmup, precision, f32, tune=False,
            opts=None):
    from tandav_amd import Engine
    if precision == "fp32":
        e = Engine(periodic=True, boxsize=1.0, softening=soft, theta_bh=0.1,
                   alpha_rel=f32["alpha"], compat=False, precision="fp32",
                   ewald_size=f32["ewald_size"],
                   ewald_accuracy=f32["ewald_accuracy"],
                   fp32_order=f32["fp32_order"],
                   fp32_criterion=f32["fp32_criterion"])
    else:
        e = Engine(periodic=True, boxsize=1.0, softening=soft, theta_bh=0.1,
                   alpha_rel=w["alpha"], compat=True, precision="fp64")
    e.set_particles(pos, m)
    e.setup_domain()
    e.domain_decomposition()
    log("%s: prepare step (BH walk, then relative walk)" % precision)
    e.compute_acceleration(prepare_step=True)
    e.autotune = None
    if tune and precision == "fp32":
        e.autotune = tune_walk(e)
    elif opts is not None and precision == "fp32":
        e.set_fp32_options(*opts)
    for i in range(warmup):
        e.domain_decomposition()
        e.compute_acceleration(prepare_step=False)
        log("%s: warmup step %d" % (precision, i))
    stage = {k: 0.0 for k in e.stage_times()}
    t0 = time.perf_counter()
    for _ in range(steps):
        e.domain_decomposition()
        e.compute_acceleration(prepare_step=False)
        for k, v in e.stage_times().items():
            stage[k] += v
        log("%s: timed step done" % precision)
    t1 = time.perf_counter()
    return e, (t1 - t0) / steps, {k: v / steps for k, v in stage.items()}


# Algorithmic HBM bytes of the memory-bound stages, per particle (and per
# tree node for the build): keys read 3 positions, write a 16-byte key; the
# sort is 8 LSD passes over the keys' high 64-bit words with a 4-byte
# payload (per pass a histogram read of the key, then key + payload read
# and written) and the permutation's copy; the reorder reads and writes 148
# payload bytes (key 16, fifteen FP64 columns, three 32-bit columns); the
# build reads key, position and mass and writes the tree-parent index, and
# per node its CoM/mass, centre/size, links and DUp.
STAGE_BYTES = {"keys": 24 + 16, "sort": 8 * (8 + 2 * (8 + 4)) + 2 * 4,
               "reorder": 2 * (16 + 15 * 8 + 3 * 4),
               "build": 16 + 24 + 8 + 4}
BUILD_NODE_BYTES = 32 + 32 + 16 + 4


def stage_roofline(stages_ms, n, nnodes):
    """HBM rate of the keys, sort, reorder and build stages against 8 TB/s
    from their algorithmic bytes (STAGE_BYTES) and HIP-event times."""
    out = {}
    for k, per in STAGE_BYTES.items():
        ms = stages_ms.get(k, 0.0)
        b = per * n + (BUILD_NODE_BYTES * nnodes if k == "build" else 0)
        if ms > 0:
            out[k] = {"bytes": b, "ms": round(ms, 4),
                      "GB_s": round(b / (ms * 1e-3) / 1e9, 1),
                      "frac_of_8TB_s": round(b / (ms * 1e-3) / 8e12, 4)}
    return out


def tune_walk(e, active_sets=()):
    """The autotuner's pick of the FP32 walk's launch options on this tree
    (untimed; tandav_amd/autotune.py)."""
    from tandav_amd.autotune import autotune
    t0 = time.perf_counter()
    r = autotune(e, active_sets=active_sets)
    log("autotune: waves %d, queue %d, split %d (%.1fs)"
        % (r["waves"], r["queue"], r["split"], time.perf_counter() - t0))
    return r


def measured_traffic(kernel, workload, f32):
    """HBM bytes per launch of the dominant kernel from the committed PMC
    measurement (FETCH_SIZE + WRITE_SIZE, each scaled by a plain copy's
    reading), when it was taken on this workload, these settings and this
    kernel source; else None."""
    import hashlib
    path = os.path.join(ROOT, "profiles", "walk32_%s_traffic.json" % workload)
    src = os.path.join(ROOT, "tandav_amd", "csrc", "tdv_walk32.hip")
    try:
        with open(path) as f:
            d = json.load(f)
        with open(src, "rb") as f:
            sha = hashlib.sha256(f.read()).hexdigest()
    except OSError:
        return None
    same = all(abs(d["fp32"][k] - f32[k]) < 1e-12 for k in d["fp32"])
    if (d["kernel"] != kernel or d["workload"] != workload or not same
            or d["walk32_source_sha256"] != sha):
        return None
    return d["traffic_bytes"]


def errors(a, ref):
    r = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
    return {"median": float(np.median(r)), "p99": float(np.percentile(r, 99)),
            "max": float(r.max())}


def cpu_baseline(pos, m, soft, w, sample, gpu_last, gpu_cost):
    """Oracle on the host cores: decomposition + build on every particle,
    then the steady-state walks (relative + Ewald) on a random sample of
    particles with the same previous-step |a| the GPU step used, so the
    sample does exactly the GPU's work (its Cost is checked against the
    GPU's) and scales to the whole step."""
    import oracle
    n = len(m)
    host = host_cpu()
    nthreads = host["usable_cores"]
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        nthreads = min(nthreads, int(os.environ["OMP_NUM_THREADS"]))
    log("cpu: %d threads (nproc %s, usable %d, %s)" % (
        nthreads, host["nproc"], host["usable_cores"], host["model"]))
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
    s.last_acc_mag[:] = gpu_last          # same PH order as the GPU arrays
    t3 = time.perf_counter()
    s.tree_acceleration(use_bh=False)
    t4 = time.perf_counter()
    log("cpu: relative+Ewald walk on %d particles %.2fs" % (sample, t4 - t3))
    same_cost = bool(np.array_equal(s.cost[pick], gpu_cost[pick]))
    t_step = (t1 - t0) + (t2 - t1) + (t4 - t3) * n / sample
    return {"value": n / t_step, "unit": "force-evals/s", "cores": nthreads,
            "kind": "port", "host": host,
            "ntarget": "NTarget = 1 (level-2 top nodes, the decomposition the "
                       "reference makes with 1 thread or with >= 65 threads, "
                       "domain.c:153-156; the GPU runs the same one)",
            "sample": "oracle (C99+OpenMP restatement of the reference path, "
                      "FP64 walk): decomposition %.2fs + tree build %.2fs on "
                      "all %d particles, steady-state relative+Ewald walk on "
                      "%d random particles %.2fs scaled by N/sample; sample "
                      "Cost identical to the GPU FP64 walk's: %s"
                      % (t1 - t0, t2 - t1, n, sample, t4 - t3, same_cost)}


# alpha 0.003: the clustered box's halo cores need a tighter criterion than
# the near-uniform boxes for p99 <= 2e-3 against the exact Ewald sums (2^20
# particles: p99 2.0e-3 at 0.005, 8e-4 at 0.0025; 2^23: 1.9e-3 at 0.004 on
# 512 samples, too close to the bar)
C4 = dict(n=1 << 23, nhalo=256, seed=11, alpha=0.002,
          desc="clustered periodic box: 256 "
          "Hernquist halos (dN/dM ~ M^-1.9 over 3 decades, r_s ~ M^(1/3)) + "
          "20% uniform background, 2^23 = 8,388,608 particles, 8 time bins")


def walk32_kernel(waves, queue, split, nactive):
    """Name of the kernel an eager FP32 walk over nactive particles runs:
    the work-queue kernel, the one-workgroup-per-item kernel when the walk
    splits (parts per group as walk32_parts: up to 32768 items from 4096
    groups up, 16384 below, at most 64 parts or `split`), else the plain
    grid; with the register budget."""
    ngroups = (nactive + 63) // 64
    kmax = split if split >= 2 else 64
    budget = 32768 if ngroups >= 4096 else 16384
    parts = 1 if split == 1 or ngroups <= 0 else max(1, min(budget // ngroups,
                                                             kmax))
    if queue:
        return "k_walk32_queue"
    base = "k_walk32_items" if parts > 1 else "k_walk32"
    return base + {0: "", 6: "_w6", 8: "_w8"}[waves]


def time_bins(e, n):
    """Time bins from |a| (the shortest steps for the largest
    accelerations): the fastest 0.05% in bin 0, fractions growing
    geometrically to everything by bin 7; returns (bins, bin-0 members)."""
    amag = np.linalg.norm(e.get("acc")["acc"], axis=1)
    order = np.argsort(-amag, kind="stable")
    frac = 0.0005 * (2000.0 ** (np.arange(8) / 7.0))
    bins = np.empty(n, np.int32)
    lo = 0
    for b in range(8):
        hi = n if b == 7 else int(round(frac[b] * n))
        bins[order[lo:hi]] = b
        lo = hi
    return bins, np.flatnonzero(bins == 0).astype(np.int32)


def bench_c4(args, f32):
    """The clustered box with hierarchical time bins: one full step (every
    particle active), then one substep per bin level with the active set of
    bins <= b (Make_Active_Particle_List, timestep.c:280-299), the tree
    kept.  Each substep runs the FP32 walk eagerly and as a replay of the
    captured walk graph (safe_last_accel + walk records + the walk over
    the active count held on the device, one launch for any active set)."""
    from tandav_amd import Engine, ics
    n = args.c4_n or C4["n"]
    log("generating clustered box, N = %d" % n)
    pos, _, m = ics.clustered_box(n, nhalo=C4["nhalo"], seed=C4["seed"])
    soft = 1.0 / (40.0 * n ** (1.0 / 3.0))
    e = Engine(periodic=True, boxsize=1.0, softening=soft, theta_bh=0.1,
               alpha_rel=f32["alpha"], compat=False, precision="fp32",
               ewald_size=f32["ewald_size"],
               ewald_accuracy=f32["ewald_accuracy"],
               fp32_order=f32["fp32_order"],
               fp32_criterion=f32["fp32_criterion"], fp32_queue=1)
    e.set_particles(pos, m)
    e.setup_domain()
    e.domain_decomposition()
    log("c4: prepare step")
    e.compute_acceleration(prepare_step=True)
    tune = None
    if not args.no_autotune:
        tune = tune_walk(e, active_sets=[time_bins(e, n)[1]])
    for i in range(args.warmup):
        e.domain_decomposition()
        e.compute_acceleration(prepare_step=False)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e.domain_decomposition()
        e.compute_acceleration(prepare_step=False)
        log("c4: timed full step done")
    t_step = (time.perf_counter() - t0) / args.steps
    stage = e.stage_times()
    newton, ewald = e.interaction_counts()
    acc = e.get("acc")["acc"]
    g = np.random.Generator(np.random.Philox(5))
    xidx = np.sort(g.choice(n, min(n, args.exact_sample), replace=False))
    t0 = time.perf_counter()
    exact = e.forcetest_exact(xidx)
    log("c4: exact-Ewald direct sum on %d particles %.1fs"
        % (len(xidx), time.perf_counter() - t0))
    err = errors(acc[xidx], exact)
    bins = time_bins(e, n)[0]
    # the graph needs one eager walk on this tree; capture, then substeps
    e.compute_acceleration(prepare_step=False)
    e.capture_walk()
    subs = []
    for b in range(8):
        act = np.flatnonzero(bins <= b).astype(np.int32)
        e.set_active(act)
        t0 = time.perf_counter()
        e.compute_acceleration(prepare_step=False)
        t_eager = time.perf_counter() - t0
        walk_eager = e.stage_times()["walk"]
        e.set_active(act)
        t0 = time.perf_counter()
        e.launch_walk()
        t_graph = time.perf_counter() - t0
        walk_graph = e.stage_times()["walk"]
        subs.append({"bin": b, "active_fraction": len(act) / n,
                     "n_active": int(len(act)), "groups": (len(act) + 63) // 64,
                     "eager_ms": t_eager * 1e3, "eager_walk_ms": walk_eager,
                     "graph_ms": t_graph * 1e3, "graph_walk_ms": walk_graph})
        if b in (0, 4):
            # the sparse sets' error: the halo cores need the tighter alpha
            acc_b = e.get("acc")["acc"]
            pick = np.sort(g.choice(act, min(len(act), 256), replace=False))
            subs[-1]["err_vs_exact_ewald"] = errors(acc_b[pick],
                                                    e.forcetest_exact(pick))
        log("c4: bins <= %d, %.4f%% active: eager %.2f ms, graph %.2f ms" % (
            b, 100.0 * len(act) / n, t_eager * 1e3, t_graph * 1e3))
    # the substep graph: records of the nodes and the active particles only
    # (the particles that a substep drifts), quadrupoles kept
    e.capture_substep_walk()
    for b, s in enumerate(subs):
        act = np.flatnonzero(bins <= b).astype(np.int32)
        e.set_active(act)
        e.launch_walk()                     # warm
        e.set_active(act)
        t0 = time.perf_counter()
        e.launch_walk()
        s["substep_graph_ms"] = (time.perf_counter() - t0) * 1e3
        s["substep_graph_walk_ms"] = e.stage_times()["walk"]
        log("c4: bins <= %d: substep graph %.2f ms" % (b,
                                                     s["substep_graph_ms"]))
    line = {
        "metric": "tree-force evals/s per step (FP32 group walk), clustered "
                  "box, and substep time vs active fraction",
        "value": n / t_step, "unit": "force-evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_step * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": C4["desc"] if n == C4["n"] else
                   C4["desc"].replace("2^23 = 8,388,608", str(n)),
                   "n_particles": n, "softening": soft,
                   # the schedule the timed walks ran (the tuner's pick)
                   "fp32": dict(f32, fp32_queue=int(e.params.fp32_queue),
                                fp32_waves=int(e.params.fp32_waves),
                                fp32_split=int(e.params.fp32_split)),
                   "error_sample_exact": len(xidx),
                   "parallelism": "single GPU"},
        "full_step": {"stages_ms": {k: round(v, 4) for k, v in stage.items()},
                      "stage_roofline": stage_roofline(stage, n,
                                                       e.num_nodes()),
                      "interactions_per_particle": (newton + ewald) / n,
                      "err_vs_exact_ewald": err},
        "substeps": subs,
        "autotune": tune,
    }
    print(json.dumps(line), flush=True)


# C5: the 512^3 box of the Zel'dovich family (P(k) ~ k^-2.5, seed 7) at rms
# |psi| = 0.5 grid spacing, one full step at a = 0.1 of a flat LCDM run
# (Omega_m 0.3): keys, sort, build, FP32 walk + Ewald, comoving kick-drift.
# Units: box 1, total mass 1, G 1, so H0 = sqrt(8 pi / (3 Omega_m)).
C5 = dict(ngrid=512, seed=7, rms=0.5, a=0.1, dlna=0.01, omega_m=0.3,
          omega_l=0.7,
          desc="periodic box, Zel'dovich P(k) ~ k^-2.5 (seed 7) at rms |psi| "
          "0.5 grid spacing, 512^3 = 134,217,728 particles, full step at "
          "a = 0.1 (flat LCDM, Omega_m 0.3, dln a 0.01)")


def bench_domain(args, f32):
    """Domain_Decomposition with NTarget > 1 (domain.c:67-145) on the
    clustered box, P.Cost from one FP32 walk: the whole decomposition
    (keys, sort, reorder, the bunch loop, top nodes) per target count, the
    bunch loop against the oracle's on the same keys and costs (its time is
    the CPU baseline), and the single-target decomposition beside it."""
    import oracle
    from tandav_amd import Engine, ics
    n = args.c4_n or C4["n"]
    log("domain: generating clustered box, N = %d" % n)
    pos, _, m = ics.clustered_box(n, nhalo=C4["nhalo"], seed=C4["seed"])
    soft = 1.0 / (40.0 * n ** (1.0 / 3.0))
    e = Engine(periodic=True, boxsize=1.0, softening=soft, theta_bh=0.1,
               alpha_rel=C4["alpha"], compat=False, precision="fp32",
               ewald_size=f32["ewald_size"],
               ewald_accuracy=f32["ewald_accuracy"],
               fp32_order=f32["fp32_order"],
               fp32_criterion=f32["fp32_criterion"])
    e.set_particles(pos, m)
    e.setup_domain()
    e.domain_decomposition()
    e.compute_acceleration(prepare_step=True)      # P.Cost of every particle
    log("domain: costs from one walk")
    e.domain_decomposition()
    t0 = time.perf_counter()
    e.domain_decomposition()
    t_one = time.perf_counter() - t0
    one = e.stage_times()
    rows = []
    for nt in [int(v) for v in args.ntargets.split(",")]:
        e.domain_decomposition_targets(nt)          # warm, same input
        t0 = time.perf_counter()
        tasks, passes = e.domain_decomposition_targets(nt)
        t = time.perf_counter() - t0
        st = e.stage_times()
        keys = e.keys(reversed_=False)[:, 1].copy()
        cost = e.get("cost")["cost"]
        t0 = time.perf_counter()
        rec, passes_o = oracle.domain_targets(keys, cost, nt)
        t_cpu = time.perf_counter() - t0
        top = e.top_nodes()
        same = (passes_o == passes and len(rec) == len(tasks)
                and bool(np.array_equal(-(rec["Target"] + 1), tasks))
                and bool(np.array_equal(rec["Key"], top["Key"]))
                and bool(np.array_equal(rec["First_Part"], top["First_Part"])))
        w = np.maximum(1, cost).astype(np.int64)
        bc = np.add.reduceat(w, top["First_Part"][::-1])[::-1]
        tc = np.bincount(tasks, weights=bc, minlength=nt)
        rows.append({"ntarget": nt, "ms": t * 1e3,
                     "stages_ms": {k: round(st[k], 4) for k in
                                   ("keys", "sort", "reorder", "top_nodes")},
                     "passes": passes, "top_nodes": len(tasks),
                     "max_cost_imbalance": float(np.max(np.abs(tc - tc.mean())
                                                        / tc.mean())),
                     "oracle_loop_ms": t_cpu * 1e3, "matches_oracle": same})
        log("domain: ntarget %d: %.2f ms (bunch loop %.2f ms), %d passes, "
            "%d top nodes, oracle loop %.1f ms, same %s" % (
                nt, t * 1e3, st["top_nodes"], passes, len(tasks),
                t_cpu * 1e3, same))
    head = next((r for r in rows if r["ntarget"] == 16), rows[-1])
    line = {
        "metric": "particles/s through Domain_Decomposition with NTarget "
                  "targets (clustered box)",
        "value": n / (head["ms"] * 1e-3), "unit": "particles/s",
        "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": head["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64/f32", "data": "synthetic",
        "config": {"workload": C4["desc"] if n == C4["n"] else
                   C4["desc"].replace("2^23 = 8,388,608", str(n)),
                   "n_particles": n, "ntarget": head["ntarget"],
                   "cost": "P.Cost of one FP32 walk",
                   "parallelism": "single GPU"},
        "ntargets": rows,
        "single_target": {"ms": t_one * 1e3,
                          "stages_ms": {k: round(one[k], 4) for k in
                                        ("keys", "sort", "reorder",
                                         "top_nodes")}},
        "cpu_baseline": {"value": n / (head["oracle_loop_ms"] * 1e-3),
                         "unit": "particles/s", "cores": 1, "kind": "port",
                         "sample": "the oracle's bunch loop (fills, "
                                   "distribute, refinement) on the same "
                                   "sorted keys and costs, keys and sort "
                                   "excluded"},
    }
    print(json.dumps(line), flush=True)


def bench_sort(args):
    """Domain_Decomposition's keys, sort and reorder (domain.c:67-145,
    sort.c:253-341, peano.c:50-82) on ngrid^3 uniform random positions in
    a periodic box (default 512^3 = 134,217,728): each stage's time over
    --steps decompositions after --warmup, and the sort's algorithmic HBM
    rate (264 B per particle: 8 passes of 8 B read for the histogram and
    12 B read + 12 B written by the scatter, the first reading 128-bit keys
    for their high words)."""
    from tandav_amd import Engine
    ng = args.sort_ngrid
    n = ng ** 3
    g = np.random.Generator(np.random.Philox(1))
    pos = g.random((n, 3))
    m = np.full(n, 1.0 / n)
    e = Engine(periodic=True, boxsize=1.0, softening=1.0 / (40 * ng),
               alpha_rel=FP32["alpha"], compat=False, precision="fp32")
    e.set_particles(pos, m)
    e.setup_domain()
    for _ in range(args.warmup):
        e.domain_decomposition()
    rows = []
    for _ in range(args.steps):
        e.domain_decomposition()
        st = e.stage_times()
        rows.append({k: float(st[k]) for k in ("keys", "sort", "reorder")})
        log("sort: %d keys: keys %.3f ms, sort %.3f ms, reorder %.3f ms"
            % (n, st["keys"], st["sort"], st["reorder"]))
    sort_ms = float(np.mean([r["sort"] for r in rows]))
    line = {
        "metric": "keys sorted/s (PH keys, LSD radix sort of the high words, "
                  "inside the domain decomposition)",
        "value": n / (sort_ms * 1e-3), "unit": "keys/s", "n_gpus": 1,
        "steps": len(rows), "warmup": args.warmup, "ms_per_step": sort_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": "%d^3 = %d uniform random positions, "
                   "periodic unit box" % (ng, n), "n_particles": n,
                   "parallelism": "single GPU"},
        "stages_ms": rows,
        "sort_algorithmic_GB_s": 264.0 * n / (sort_ms * 1e-3) * 1e-9,
    }
    print(json.dumps(line), flush=True)


def bench_snapshot(args):
    """Write_Snapshot and Read_Snapshot (IO/write_snapshot.c,
    IO/read_snapshot.c) of the 128^3 box through the C ABI, in a local
    temporary directory: particles/s each way and the file's MB/s, and the
    read-back columns checked bit for bit against the written ones."""
    import tempfile
    from tandav_amd import Engine
    pos, vel, m, soft, w = make_workload("c2")
    n = len(m)
    e = Engine(periodic=True, boxsize=1.0, softening=soft)
    e.set_particles(pos, m, vel=vel, ids=np.arange(n))
    rows = []
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "snap_000")
        for rep in range(2):
            t0 = time.perf_counter()
            e.write_snapshot(path, npart=[0, n, 0, 0, 0, 0], boxsize=1.0)
            t_w = time.perf_counter() - t0
            size = os.path.getsize(path)
            e2 = Engine(periodic=True, boxsize=1.0, softening=soft)
            t0 = time.perf_counter()
            e2.read_snapshot(path)
            t_r = time.perf_counter() - t0
            a, b = e.get("pos", "vel", "mass", "id"), e2.get("pos", "vel",
                                                           "mass", "id")
            same = all(np.array_equal(a[k], b[k]) for k in a)
            e2.close()
            rows.append({"write_s": t_w, "read_s": t_r, "bytes": size,
                         "same": same})
            log("snapshot: %d particles, %.1f MB: write %.3f s, read %.3f s,"
                " same %s" % (n, size / 1e6, t_w, t_r, same))
    best = min(rows, key=lambda r: r["write_s"] + r["read_s"])
    line = {
        "metric": "particles/s through Write_Snapshot + Read_Snapshot "
                  "(Gadget format 2, one file)",
        "value": n / (best["write_s"] + best["read_s"]),
        "unit": "particles/s", "n_gpus": 1, "steps": len(rows), "warmup": 0,
        "ms_per_step": (best["write_s"] + best["read_s"]) * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": w["desc"], "n_particles": n,
                   "blocks": "HEAD POS VEL ID MASS ACC GACC",
                   "parallelism": "single GPU, host file I/O"},
        "write_MB_s": best["bytes"] / best["write_s"] / 1e6,
        "read_MB_s": best["bytes"] / best["read_s"] / 1e6,
        "bytes": best["bytes"], "runs": rows,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def bench_c5(args, f32):
    """One full step of the large box with every particle active, device
    resident: PH keys, radix sort, reorder, top nodes, build, the FP32 walk
    with its Ewald correction, then the comoving kick (the step's two half
    kicks) and drift.  The step is run twice from the same positions: the
    first, untimed, is scored against the oracles (the table direct sum of
    Gravity_Forcetest on --c5-sample particles and the exact Ewald sums on
    --exact-sample), the second is timed with the kick-drift."""
    import torch
    from tandav_amd import Engine, ics
    from tandav_amd.comoving import ComovingFactors, e_hubble
    ng = args.c5_ngrid or C5["ngrid"]
    n = ng ** 3
    a0 = C5["a"]
    a1 = a0 * np.exp(C5["dlna"])
    om, ol = C5["omega_m"], C5["omega_l"]
    h0 = np.sqrt(8.0 * np.pi / (3.0 * om))
    hub = h0 * e_hubble(a0, om + ol, om, ol)
    om_a = om / (a0 ** 3 * (hub / h0) ** 2)
    vel_factor = a0 * hub * om_a ** 0.55          # comoving a H f
    log("c5: generating %d^3 Zel'dovich box" % ng)
    t0 = time.perf_counter()
    pos, vel, m = ics.zeldovich_box(ng, seed=C5["seed"], rms_frac=C5["rms"],
                                    vel_factor=vel_factor)
    log("c5: initial conditions %.1fs" % (time.perf_counter() - t0))
    fac = ComovingFactors(0.9 * a0, 1.1 * a1, hubble=h0, omega_matter=om,
                          omega_lambda=ol)
    kick, drift = fac.kick(a0, a1), fac.drift(a0, a1)
    soft = 1.0 / (40.0 * ng)
    free0 = torch.cuda.mem_get_info()[0]
    e = Engine(periodic=True, boxsize=1.0, softening=soft, theta_bh=0.1,
               alpha_rel=f32["alpha"], compat=False, precision="fp32",
               ewald_size=f32["ewald_size"],
               ewald_accuracy=f32["ewald_accuracy"],
               fp32_order=f32["fp32_order"],
               fp32_criterion=f32["fp32_criterion"])
    e.set_particles(pos, m, vel=vel)
    del pos, vel
    e.setup_domain()
    e.domain_decomposition()
    log("c5: prepare step")
    t0 = time.perf_counter()
    e.compute_acceleration(prepare_step=True)
    log("c5: prepare step %.1fs" % (time.perf_counter() - t0))
    tune = None if args.no_autotune else tune_walk(e)
    e.domain_decomposition()
    e.compute_acceleration(prepare_step=False)
    log("c5: scored step done")
    acc = e.get("acc")["acc"]
    g = np.random.Generator(np.random.Philox(5))
    idx = np.sort(g.choice(n, min(n, args.c5_sample), replace=False))
    t0 = time.perf_counter()
    ref = e.forcetest(idx)
    log("c5: table direct sum on %d particles %.1fs"
        % (len(idx), time.perf_counter() - t0))
    xidx = idx[np.sort(np.random.Generator(np.random.Philox(6)).choice(
        len(idx), min(len(idx), args.c5_exact_sample), replace=False))]
    t0 = time.perf_counter()
    exact = e.forcetest_exact(xidx)
    log("c5: exact Ewald sums on %d particles %.1fs"
        % (len(xidx), time.perf_counter() - t0))
    err_table = errors(acc[idx], ref)
    err_exact = errors(acc[xidx], exact)
    pos_idx = {int(j): k for k, j in enumerate(idx)}
    err_table_vs_exact = errors(ref[[pos_idx[int(j)] for j in xidx]], exact)
    del acc
    log("c5: timed step")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.domain_decomposition()
    e.compute_acceleration(prepare_step=False)
    e.kick_drift(kick, drift)
    t_step = time.perf_counter() - t0
    stage = e.stage_times()
    newton, ewald = e.interaction_counts()
    kinds = e.interaction_kinds()
    free1 = torch.cuda.mem_get_info()[0]
    walk_s = stage["walk"] / 1e3
    flops = walk32_flops(kinds)
    line = {
        "metric": "tree-force evals/s per full step (FP32 group walk), "
                  "512^3 box, at p99 |da|/|a| <= 2e-3",
        "value": n / t_step, "unit": "force-evals/s",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": t_step * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": C5["desc"] if ng == C5["ngrid"] else
                   C5["desc"].replace("512^3 = 134,217,728",
                                      "%d^3 = %d" % (ng, n)),
                   "n_particles": n, "softening": soft,
                   "fp32": f32, "kick": kick, "drift": drift,
                   "parallelism": "single GPU"},
        "stages_ms": {k: round(v, 3) for k, v in stage.items()},
        "stage_roofline": stage_roofline(stage, n, e.num_nodes()),
        "walk_force_evals_per_s": n / walk_s,
        "interactions_per_s": (newton + ewald) / walk_s,
        "interactions_per_particle": (newton + ewald) / n,
        "walk_roofline": {"bound": "valu",
                          "kernel": walk32_kernel(e.params.fp32_waves,
                                                  e.params.fp32_queue,
                                                  e.params.fp32_split, n),
                          "achieved": flops / walk_s / 1e12,
                          "peak": 157.3, "unit": "TFLOP/s",
                          "frac": flops / walk_s / 1e12 / 157.3,
                          "counting": COUNTING},
        "hbm_bytes_per_particle": (free0 - free1) / n,
        "autotune": tune,
        "accuracy": {"vs_table_direct_sum": dict(err_table, n=len(idx)),
                     "vs_exact_ewald": dict(err_exact, n=len(xidx)),
                     "table_direct_sum_vs_exact": err_table_vs_exact},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3",
                    choices=sorted(WORKLOADS) + ["c4", "c5", "domain",
                                                 "snapshot", "sort"])
    ap.add_argument("--sort-ngrid", type=int, default=512,
                    help="the sort workload's ngrid^3 particles")
    ap.add_argument("--ntargets", default="2,8,16,64",
                    help="target counts of the domain workload")
    ap.add_argument("--c4-n", type=int, default=0,
                    help="particles of the clustered box (default 2^23)")
    ap.add_argument("--c5-ngrid", type=int, default=0,
                    help="grid of the large box (default 512)")
    ap.add_argument("--c5-sample", type=int, default=16384,
                    help="particles scored against the table direct sum")
    ap.add_argument("--c5-exact-sample", type=int, default=16,
                    help="particles scored against the exact Ewald sums "
                         "(about 4 s each at 512^3)")
    ap.add_argument("--fp32-alpha", type=float, default=None,
                    help="FP32 walk's relative opening criterion (default "
                         "%g, %g for the clustered box)"
                         % (FP32["alpha"], C4["alpha"]))
    ap.add_argument("--ewald-size", type=float, default=FP32["ewald_size"])
    ap.add_argument("--ewald-accuracy", type=float,
                    default=FP32["ewald_accuracy"])
    ap.add_argument("--fp32-criterion", type=int,
                    default=FP32["fp32_criterion"])
    ap.add_argument("--error-sample", type=int, default=8192)
    ap.add_argument("--exact-sample", type=int, default=512)
    ap.add_argument("--cpu-sample", type=int, default=32768,
                    help="particles whose FP64 walks the CPU baseline times")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the FP32 walk's default launch options "
                         "instead of the autotuner's pick")
    ap.add_argument("--fp32-options", default=None, metavar="W,Q,S",
                    help="FP32 walk's launch options (waves, queue, split) "
                         "instead of the autotuner's pick, e.g. 6,0,0 for a "
                         "profile of the kernel the autotuner picks")
    ap.add_argument("--only", choices=["fp32", "fp64"], default=None)
    ap.add_argument("--alphas", default=None,
                    help="comma-separated further opening criteria for the "
                         "FP32 walk, one timed step each, scored like the "
                         "headline (default 0.005,0.0025 on the 256^3 box)")
    ap.add_argument("--fp64-steps", type=int, default=2,
                    help="timed regular steps of the FP64 walk after its "
                         "prepare step")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device (one plain copy "
                         "kernel, the reference point for the HBM byte "
                         "counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        nb = args.copy_calibration << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": nb}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus %d is not "
              "supported (multi-GPU sharding is a design note only)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    alpha = args.fp32_alpha
    if alpha is None:
        alpha = C4["alpha"] if args.workload == "c4" else FP32["alpha"]
    f32 = dict(FP32, alpha=alpha, ewald_size=args.ewald_size,
               ewald_accuracy=args.ewald_accuracy,
               fp32_criterion=args.fp32_criterion)
    if args.alphas is None:
        args.alphas = "0.005,0.0025" if args.workload == "c3" else ""
    if args.workload == "c4":
        bench_c4(args, f32)
        return
    if args.workload == "c5":
        bench_c5(args, f32)
        return
    if args.workload == "domain":
        bench_domain(args, f32)
        return
    if args.workload == "snapshot":
        bench_snapshot(args)
        return
    if args.workload == "sort":
        bench_sort(args)
        return

    bench_box(args, f32)


def host_cpu():
    """nproc, the cores this process may run on, and the CPU model."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "usable_cores": usable, "model": model}


def bench_box(args, f32):
    """The Zel'dovich boxes (default: 256^3).  The FP64 walk (the drop-in:
    the reference's walk, relative criterion 0.005) runs its prepare step
    (BH walk, then relative walk) and --fp64-steps timed regular steps; the
    FP32 group walk runs its prepare step, the autotuner, --warmup and
    --steps timed regular steps at the headline opening criterion, then one
    timed step at each further criterion of --alphas.  Every walk is scored
    against the exact periodic force on --exact-sample particles and
    against Gravity_Forcetest's table sum on --error-sample."""
    log("generating %s" % args.workload)
    pos, vel, m, soft, w = make_workload(args.workload)
    n = len(m)
    log("N = %d" % n)
    g = np.random.Generator(np.random.Philox(5))
    idx = np.sort(g.choice(n, min(n, args.error_sample), replace=False))
    xidx = idx[np.sort(np.random.Generator(np.random.Philox(6)).choice(
        len(idx), min(len(idx), args.exact_sample), replace=False))]
    tpos = np.searchsorted(idx, xidx)
    res = {}
    ref = exact = table_err = None
    cpu = None

    def score(e):
        nonlocal ref, exact, table_err
        if ref is None:
            # particle order is the same PH order for every walk
            t0 = time.perf_counter()
            ref = e.forcetest(idx)
            log("direct sum with the table's Ewald on %d particles %.2fs"
                % (len(idx), time.perf_counter() - t0))
            t0 = time.perf_counter()
            exact = e.forcetest_exact(xidx)
            log("direct sum with the exact periodic force on %d particles "
                "%.2fs" % (len(xidx), time.perf_counter() - t0))
            table_err = errors(ref[tpos], exact)
        acc = e.get("acc")["acc"]
        return errors(acc[xidx], exact), errors(acc[idx], ref)

    if args.only != "fp32":
        e, t_step, stage = run_gpu(pos, m, soft, w, args.fp64_steps, 0,
                                   "fp64", f32)
        err, err_t = score(e)
        newton, ewald = e.interaction_counts()
        got = e.get("cost", "last_acc_mag")
        res["fp64"] = dict(t_step=t_step, stage=stage, err=err,
                           err_table=err_t, newton=newton, ewald=ewald,
                           steps=args.fp64_steps, nnodes=e.num_nodes())
        log("fp64: %.1f ms/step, err vs exact %s" % (t_step * 1e3, err))
        e.close()
        if not args.no_cpu_baseline:
            cpu = cpu_baseline(pos, m, soft, w, min(args.cpu_sample, n),
                               got["last_acc_mag"], got["cost"])
        del got
    tune = None
    if args.only != "fp64":
        fixed = (tuple(int(x) for x in args.fp32_options.split(","))
                 if args.fp32_options else None)
        e, t_step, stage = run_gpu(pos, m, soft, w, args.steps, args.warmup,
                                   "fp32", f32,
                                   tune=not args.no_autotune and not fixed,
                                   opts=fixed)
        err, err_t = score(e)
        newton, ewald = e.interaction_counts()
        kinds = e.interaction_kinds()
        tune = e.autotune
        pe = e.params
        opts = (int(pe.fp32_waves), int(pe.fp32_queue), int(pe.fp32_split))
        res["fp32"] = dict(t_step=t_step, stage=stage, err=err,
                           err_table=err_t, newton=newton, ewald=ewald,
                           kinds=kinds,
                           steps=args.steps, nnodes=e.num_nodes(),
                           kernel=walk32_kernel(*opts, n), opts=opts)
        log("fp32: %.1f ms/step, err vs exact %s" % (t_step * 1e3, err))
        e.close()
    scan = []
    alphas = [float(x) for x in args.alphas.split(",") if x]
    for a in (alphas if args.only != "fp64" else []):
        fa = dict(f32, alpha=a)
        e, t_step, stage = run_gpu(pos, m, soft, w, 1, 0, "fp32", fa,
                                   opts=res["fp32"]["opts"])
        err, err_t = score(e)
        newton, ewald = e.interaction_counts()
        scan.append({"alpha": a, "ms_per_step": t_step * 1e3,
                     "force_evals_per_s": n / t_step,
                     "walk_ms": stage["walk"],
                     "interactions_per_particle": (newton + ewald) / n,
                     "err_vs_exact_ewald": err, "err_vs_table_ewald": err_t,
                     "meets_bar": err["p99"] <= 2e-3})
        log("fp32 alpha %g: %.1f ms/step, err vs exact %s" % (
            a, t_step * 1e3, err))
        e.close()

    head = res["fp32"] if "fp32" in res else res["fp64"]
    if "fp32" in res:
        kern, t_k = head["kernel"], head["stage"]["walk"]
        flops = walk32_flops(head["kinds"])
        peak, dtype = PEAK["f32"], "f32"
    else:
        walk_ms, ewald_ms = head["stage"]["walk"], head["stage"]["ewald"]
        if walk_ms >= ewald_ms:
            kern, t_k = "k_walk64", walk_ms
            flops = head["newton"] * FLOPS_PER_INTERACTION
        else:
            kern, t_k = "k_walk64_ewald", ewald_ms
            flops = head["ewald"] * FLOPS_PER_EWALD
        peak, dtype = PEAK["f64"], "f64"
    achieved = flops / (t_k * 1e-3)
    roof = {"bound": "valu", "kernel": kern, "achieved": achieved / 1e12,
            "peak": peak / 1e12, "unit": "TFLOP/s", "frac": achieved / peak,
            "traffic": measured_traffic(kern, args.workload, f32),
            "avg_launch_ms": t_k,
            "flops_per_launch": flops,
            "counting": (COUNTING if "fp32" in res else
                         "20 flops per monopole interaction, 60 per Ewald "
                         "correction; interactions counted on the device")}

    def summary(r):
        walk_s = (r["stage"]["walk"] + r["stage"]["ewald"]) * 1e-3
        return {"force_evals_per_s": n / r["t_step"],
                "ms_per_step": r["t_step"] * 1e3, "timed_steps": r["steps"],
                "walk_force_evals_per_s": n / walk_s,
                "interactions_per_s": (r["newton"] + r["ewald"]) / walk_s,
                "newton_interactions": r["newton"],
                "ewald_interactions": r["ewald"],
                "stages_ms": {k: round(v, 4) for k, v in r["stage"].items()},
                "stage_roofline": stage_roofline(r["stage"], n, r["nnodes"]),
                "err_vs_exact_ewald": r["err"],
                "err_vs_table_ewald": r["err_table"]}

    line = {
        "metric": "tree-force evals/s per step, FP32 vs FP64 walk, at p99 "
                  "|da|/|a| <= 2e-3",
        "value": n / head["t_step"],
        "unit": "force-evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["t_step"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype, "data": "synthetic",
        "config": {"workload": w["desc"], "n_particles": n,
                   "softening": soft, "fp64_alpha_rel": w["alpha"],
                   "fp32": dict(f32), "fp32_alphas_scanned": alphas,
                   "fp64_steps": "prepare step (untimed) + %d timed"
                                 % args.fp64_steps,
                   "error_sample_exact": len(xidx),
                   "error_sample_table": len(idx),
                   "parallelism": "single GPU"},
        "accuracy": {
            "oracle": "FP64 direct sum over all particles with the exact "
                      "periodic force (Ewald lattice sum per pair, no table)",
            "p99_bar": 2e-3,
            "table_oracle_vs_exact": table_err},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    for prec, r in res.items():
        line[prec] = summary(r)
        line["accuracy"][prec + "_p99"] = r["err"]["p99"]
        line["accuracy"][prec + "_meets_bar"] = r["err"]["p99"] <= 2e-3
    if tune:
        line["fp32"]["autotune"] = tune
    if "fp32" in res and "fp64" in res:
        line["fp32_over_fp64"] = res["fp64"]["t_step"] / res["fp32"]["t_step"]
    if scan:
        line["fp32_alpha_scan"] = scan
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

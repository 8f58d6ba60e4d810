"""A fresh program running Tandav's calling sequence through the binding
with the FP32 group walk (TANDAV_WALK=fp32 set by the caller before any GPU
call): a prepare step and a regular step, then |da|/|a| of the binding's
P.Acc against FP64 direct summation (the exact periodic force on periodic
runs) on 512 sampled particles, printed as one JSON line.
Usage: python fp32_step.py periodic(0|1)"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tandav_amd import Engine, ics  # noqa: E402
import host as glue_host  # noqa: E402


def main():
    periodic = sys.argv[1] == "1"
    assert os.environ.get("TANDAV_WALK") == "fp32"
    if periodic:
        pos, vel, m = ics.zeldovich_box(32, seed=7)
        soft = 1 / 1280.
    else:
        pos, vel, m = ics.hernquist_sphere(65536, seed=1)
        soft = 0.01
    e = Engine(periodic=periodic, softening=soft)
    e.set_particles(pos, m, vel=vel)
    e.setup_domain()
    e.domain_decomposition()
    top = e.top_nodes()
    size, origin, _ = e.domain
    h = glue_host.Host(periodic)
    h.init(pos, vel, m, boxsize=1.0, softening=soft, theta_bh=0.1,
           alpha_rel=0.005, threads=4)
    for prepare in (1, 0):
        h.decomposition(size, origin, top)
        h.L.host_compute_acceleration(prepare)
    assert np.array_equal(h.vec("Pos"), e.get("pos")["pos"])
    acc = h.vec("Acc")
    g = np.random.Generator(np.random.Philox(5))
    idx = np.sort(g.choice(len(m), 512, replace=False))
    ref = e.forcetest_exact(idx)
    r = np.linalg.norm(acc[idx] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print(json.dumps({"periodic": periodic, "median": float(np.median(r)),
                      "p99": float(np.percentile(r, 99)),
                      "cost_min": float(h.column("Cost").min())}),
          flush=True)


if __name__ == "__main__":
    main()

"""Benchmark of the FD time-stepping hot path on one MI355X.

Workload (one "step" = one leapfrog time step of the whole grid):
  512^3 interior points, radius-4 (8th-order) Laplacian, 2nd-order leapfrog,
  20-point sponge, vp layered 1500-4500 m/s in 6 seeded layers, Ricker
  f0 = 10 Hz, dt from compute_dt_sch with CFL 0.4, FP32 fields.

The timed region runs K steps through phc_fd_run with per-launch HIP events
on the solver's stream; value = interior points x K / wall time.  The
roofline block uses the mean stencil-kernel duration from those events and
16 algorithmic bytes per point update (read pn, pnm1, vp^2dt^2; write pnp1).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "proxy-geos-hc_amd"))

METRIC = "Mpoint-updates/s per MI355X (SEM & FD); learned vs sponge at ≤1% trace err"
HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = 16
# HBM bytes per launch of the stencil kernel, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE scaled by the same counters on a copy kernel of known size
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fd_step_r4_512_traffic.txt")


def read_traffic():
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("bytes_per_launch:"):
            return int(float(line.split(":")[1]))
    return None


def cpu_baseline(setup, nsteps=3):
    """The oracle (a restatement of the reference's FD loop: two passes per
    step, FP64 combine) on this host's cores, on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    P = oracle.FdProblem(nx=setup.nx, model=setup.vp, f0=setup.f0, dt=setup.dt,
                         nsamples=setup.nsamples)
    P.run(1)  # warm the pages
    t0 = time.perf_counter()
    P.run(nsteps)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    pts = setup.nx * setup.ny * setup.nz
    return {"value": pts * nsteps / dt / 1e6, "unit": "Mpoint-updates/s", "cores": cores,
            "kind": "port",
            "sample": "%d steps of the same 512^3 layered workload (oracle/fd_oracle.c, OpenMP)" % nsteps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--arith", default="fp32", choices=["fp32", "ref"])
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU decomposition "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import phc
    if phc.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    n = args.n
    K, W = args.steps, args.warmup
    setup = phc.FdSetup(nx=n, model=("layered", 6, 20240512), f0=10.0, cfl=0.4,
                        nsamples=W + K + 1)
    arith = phc.ARITH_FP32 if args.arith == "fp32" else phc.ARITH_REF
    sol = phc.FdSolver(setup, arith=arith)
    tile_j, waves, chunk_i = sol.autotune(steps_per_trial=10)
    # warmup (untimed), then the timed region with per-launch events
    sol.set_timing(True)
    if W:
        sol.run(0, W)
    sol.synchronize()
    phc.check(phc.lib().phc_device_synchronize())
    t0 = time.perf_counter()
    sol.run(W, K)
    sol.synchronize()
    phc.check(phc.lib().phc_device_synchronize())
    wall = time.perf_counter() - t0
    kernel_ms = sol.last_kernel_ms()
    pts = n ** 3
    value = pts * K / wall / 1e6
    achieved = BYTES_PER_UPDATE * pts / (kernel_ms * 1e-3) / 1e9

    # the same K steps replayed as one captured hipGraph (no per-launch events)
    sol2 = phc.FdSolver(setup, arith=arith, tile_j=tile_j, block_waves=waves, chunk_i=chunk_i)
    if W:
        sol2.run(0, W)
    sol2.run(W, K)  # capture + first replay
    sol2.synchronize()
    t0 = time.perf_counter()
    sol2.run(W, K)  # replay of the cached graph (same steps, fresh state irrelevant to timing)
    sol2.synchronize()
    graph_ms = (time.perf_counter() - t0) / K * 1e3
    sol2.close()
    fin = sol.check_finite(sol.i2)
    # reference-rounding arithmetic (FP64 combine) on the same workload
    ref = phc.FdSolver(setup, arith=phc.ARITH_REF)
    ref.autotune(steps_per_trial=10)
    ref.set_timing(True)
    ref.run(0, min(W, 20) or 1)
    ref.run(0, min(K, 200))
    ref_kernel_ms = ref.last_kernel_ms()
    ref.close()

    traffic = read_traffic()
    # a plain 1 GiB copy: the counter calibration kernel and the practical
    # streaming bandwidth of this device
    import ctypes
    cms = ctypes.c_double()
    phc.check(phc.lib().phc_copy_bandwidth(1 << 30, 5, ctypes.byref(cms)))
    copy_gbs = 2 * (1 << 30) / (cms.value * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mpoint-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(wall / K * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded 6-layer vp model, Ricker source",
        "config": {"workload": "FD %d^3 radius-4 leapfrog, 20-pt sponge, 6 layers 1500-4500 m/s, "
                               "Ricker 10 Hz, dt CFL 0.4" % n,
                   "grid": [n, n, n], "radius": 4, "arith": args.arith,
                   "tile_j": tile_j, "block_waves": waves, "chunk_i": chunk_i},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
        "kernel_ms": round(kernel_ms, 5),
        "graph_ms_per_step": round(graph_ms, 5),
        "ref_arith_kernel_ms": round(ref_kernel_ms, 5),
        "ref_arith_mpts": round(pts / (ref_kernel_ms * 1e-3) / 1e6, 1),
        "copy_gbs": round(copy_gbs, 1),
        "finite": fin == 0,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(setup)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

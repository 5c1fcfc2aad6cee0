"""Benchmark of the FD time-stepping hot path on one MI355X.

Workload (one "step" = one leapfrog time step of the whole grid):
  512^3 interior points, radius-4 (8th-order) Laplacian, 2nd-order leapfrog,
  20-point sponge, vp layered 1500-4500 m/s in 6 seeded layers, Ricker
  f0 = 10 Hz, dt from compute_dt_sch with CFL 0.4, FP32 fields.

The timed region runs K steps through phc_fd_run with per-launch HIP events
on the solver's stream; value = interior points x K / wall time.  The
roofline block uses the mean stencil-kernel duration from those events and
16 algorithmic bytes per point update (read pn, pnm1, vp^2dt^2; write pnp1).

Beside the headline line: "batch" reports the batched multi-shot workload
(64 uneven shots, 800 steps, one persistent launch per step) and the same shots
run one at a time; "sem_p4" the SEM order-4 workload (401^3 nodes, FP32
separable kernel, and its FP32-vs-FP64 precision over 2000 steps);
"sem2d" the SEM2D step on the reference's published 2D mesh and on a large 2D
mesh; "pml" the FD step with the reference's PML; "fd_sizes" the headline FD
workload on 264^3 to 448^3 grids; "sem_element_kernels" the
element kernels (atomic and 8-colour scatter) on the reference's CPU-runnable
SEM case; "learned_boundary" the learned thin layer trained from scratch on the
GPU and compared with the 40-point sponge against a 2x-wide FP64 domain.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extras]
"""
import argparse
import ctypes
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


BATCH_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fd_batch_64shots_traffic.txt")
SEM_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sem_p4_401_traffic.txt")
SEM2D_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sem2d_p4_8001_traffic.txt")
PML_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fd_pml_r4_512_kernel_stats.txt")


def read_traffic(path=TRAFFIC_FILE):
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("bytes_per_launch:"):
            return int(float(line.split(":")[1]))
    return None


def progress(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


# Oracle against the reference on one host: oracle/time_oracle.py 8 (this
# build's oracle) beside the reference's own drivers timed on the same 8
# cores (its published FD case 150^3: 69.7 Mpoint-updates/s; its SEM case
# "50 2000": 8.17 Mnode-updates/s).  A CPU baseline here divided by the ratio
# estimates the reference's rate on the same cores.
ORACLE_OVER_REFERENCE = {
    "fd": {"oracle": 234.1, "reference": 69.7, "ratio": 3.36, "unit": "Mpoint-updates/s",
           "case": "150^3 radius 4, 20-point sponge, 8 threads, same host"},
    "sem": {"oracle": 7.85, "reference": 8.17, "ratio": 0.96, "unit": "Mnode-updates/s",
            "case": "50^3 elements of order 2, 8 threads, same host"},
}


def host_cores():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


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
    cores = host_cores()
    pts = setup.nx * setup.ny * setup.nz
    return {"value": pts * nsteps / dt / 1e6, "unit": "Mpoint-updates/s", "cores": cores,
            "kind": "port",
            "sample": "%d steps of the same 512^3 layered workload (oracle/fd_oracle.c, OpenMP)" % nsteps,
            "oracle_over_reference": ORACLE_OVER_REFERENCE["fd"]}


def cpu_baseline_sem_p4(ex=24, steps=3):
    """The SEM oracle (the reference's element loop restated: computeB V2 +
    gradPhiGradPhi per element) on every host core (OpenMP over the elements
    of one colour at a time, so the scatter does not race), on a bounded
    sample of the order-4 workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    cores = host_cores()
    P = oracle.SemProblem(ex=ex, lx=20.0 * ex, order=4, dt=2e-4, f0=8.0, nsamples=steps + 1,
                          src_point=(10.0 * ex + 2.5, 10.0 * ex + 2.5, 10.0 * ex + 2.5))
    P.set_threads(cores)
    P.run(1)
    t0 = time.perf_counter()
    P.run(steps)
    dt = time.perf_counter() - t0
    return {"value": round(P.nN * steps / dt / 1e6, 2), "unit": "Mnode-updates/s", "cores": cores, "kind": "port",
            "sample": "%d steps of a %d^3-element order-4 mesh (%d^3 nodes), oracle/sem_oracle.c, OpenMP"
                      % (steps, ex, 4 * ex + 1),
            "oracle_over_reference": ORACLE_OVER_REFERENCE["sem"]}


def cpu_baseline_sem2d(steps=200):
    """The SEM2D oracle (the reference's SEM2D element loop and paraxial edges
    restated) on one host core, on the published 2D mesh."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    O = oracle.Sem2dProblem(ex=65, order=3, lx=2000.0, dt=1e-3, f0=10.0, nsamples=steps, src_element=2178)
    t0 = time.perf_counter()
    O.run(steps)
    dt = time.perf_counter() - t0
    return {"value": round(O.nN * steps / dt / 1e6, 2), "unit": "Mnode-updates/s", "cores": 1, "kind": "port",
            "sample": "%d steps of the published 65 x 65 order-3 mesh, oracle/sem2d_oracle.c (one thread: the "
                      "2D oracle has no threaded loop)" % steps}


def bench_batch(phc, nsteps=800, nshots=64, baselines=True, shape=None):
    """Batched multi-shot FD: 64 shots, each axis drawn from U[128, 384], per-shot
    6-layer vp and Ricker f0 in U[5, 15] Hz, 20-point sponge, dt from CFL 0.4,
    nsteps steps; all shots advanced by one persistent launch per step, the run
    captured in one hipGraph.  Baselines: the same shots run one after the
    other through the single-grid path (graph-captured each), with its default
    launch shape and with a shape autotuned per shot."""
    setups = phc.multi_shot_setups(nshots=nshots, nsteps=nsteps, seed=2024)
    pts = int(sum(s.nx * s.ny * s.nz for s in setups))
    if shape:
        b = phc.FdBatch(setups, arith=phc.ARITH_FP32, tile_j=shape[0], block_waves=shape[1],
                        chunk_i=shape[2], blocks_per_cu=shape[3])
        cfg = tuple(shape)
    else:
        b = phc.FdBatch(setups, arith=phc.ARITH_FP32)
        cfg = b.autotune(steps_per_trial=3)
    b.run(nsteps, it0=0, traces=False)  # capture + warm
    b.synchronize()
    t0 = time.perf_counter()
    b.run(nsteps, it0=0, traces=False)  # replay of the captured run
    b.synchronize()
    wall = time.perf_counter() - t0
    b.set_timing(True)
    b.run(min(nsteps, 100), it0=0, traces=False)
    kernel_ms = b.last_kernel_ms()
    b.close()
    out = {"workload": "%d shots, axes U[128,384], 6-layer vp, f0 U[5,15] Hz, 20-pt sponge, "
                       "%d steps, FP32" % (nshots, nsteps),
           "points": pts, "steps": nsteps, "wall_s": round(wall, 4),
           "value": round(pts * nsteps / wall / 1e6, 1), "unit": "Mpoint-updates/s",
           "kernel_ms_per_step": round(kernel_ms, 4),
           "tile_j": cfg[0], "block_waves": cfg[1], "chunk_i": cfg[2], "blocks_per_cu": cfg[3]}
    ach = BYTES_PER_UPDATE * pts / (kernel_ms * 1e-3) / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": read_traffic(BATCH_TRAFFIC_FILE)}
    if baselines:
        t_def = t_tuned = 0.0
        for s in setups:
            for tune in (False, True):
                sol = phc.FdSolver(s, arith=phc.ARITH_FP32)
                if tune:
                    sol.autotune(steps_per_trial=3)
                sol.run(0, nsteps)  # capture + warm
                sol.synchronize()
                t0 = time.perf_counter()
                sol.run(0, nsteps)
                sol.synchronize()
                dt = time.perf_counter() - t0
                sol.close()
                if tune:
                    t_tuned += dt
                else:
                    t_def += dt
        out["one_at_a_time_default_shape"] = {"wall_s": round(t_def, 4),
                                              "value": round(pts * nsteps / t_def / 1e6, 1)}
        out["one_at_a_time_tuned_shape"] = {"wall_s": round(t_tuned, 4),
                                            "value": round(pts * nsteps / t_tuned / 1e6, 1)}
        out["speedup_vs_default_shape"] = round(t_def / wall, 3)
        out["speedup_vs_tuned_shape"] = round(t_tuned / wall, 3)
    return out


def bench_sem_p4(phc, steps=2000, fp64=True):
    """SEM order 4, 100^3 elements of 20 m (401^3 nodes), Gaussian random vp
    2000-4000 m/s, Ricker 8 Hz, dt 2e-4 s, 2000 steps: the FP32 separable
    kernel (autotuned; per-launch HIP events) and the same run in FP64."""
    import precision
    r = precision.sem_p4_study(steps=steps, fp64=fp64)
    n = r["nodes"]
    ms = r["fp32_kernel_ms_per_step"]
    ach = BYTES_PER_UPDATE * n / (ms * 1e-3) / 1e9
    r["value"] = round(n / (ms * 1e-3) / 1e6, 1)
    r["unit"] = "Mnode-updates/s"
    r["roofline"] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": read_traffic(SEM_TRAFFIC_FILE)}
    return r


# the reference's published SEM runs (2D, 65 x 65 elements of order 3, about
# 999 steps): best published rates in Mnode-updates/s
PUBLISHED_SEM2D = {"H100 (Kokkos)": 237.0, "MI100 (Kokkos)": 39.1, "Grace 72 threads (OpenMP)": 228.0}


def bench_sem2d(phc, steps=1000, big_ex=2000, big_order=4, big_steps=200):
    """SEM2D (the reference's SEM2D build): its published configuration, 65 x 65
    elements of order 3 (38,416 nodes), c = 1500 m/s, dt = 1e-3 s, Ricker 10 Hz,
    1000 steps, graph-captured and autotuned; and a large 2D mesh
    (big_ex^2 elements of order big_order) for the HBM roofline."""
    out = {}
    G = phc.Sem2dSolver(ex=65, order=3, lx=2000.0, dt=1e-3, f0=10.0, nsamples=steps + 1)
    chunk = G.autotune(steps_per_trial=20)
    G.run(steps)  # capture + warm
    G.synchronize()
    G.it = 0
    t0 = time.perf_counter()
    G.run(steps)  # replay of the captured run (samples 0..steps-1 again)
    G.synchronize()
    wall = time.perf_counter() - t0
    G.set_timing(True)
    G.it = 0
    G.run(steps)
    kms = G.last_kernel_ms()
    G.close()
    nodes = 196 * 196
    v = nodes * steps / wall / 1e6
    out["published_config"] = {
        "workload": "SEM2D 65 x 65 elements of order 3 (38,416 nodes), %d steps" % steps,
        "wall_s": round(wall, 5), "value": round(v, 1), "unit": "Mnode-updates/s",
        "us_per_step": round(wall / steps * 1e6, 2), "kernel_us": round(kms * 1e3, 2), "chunk": chunk,
        "published_best": PUBLISHED_SEM2D,
        "vs_published_best": round(v / max(PUBLISHED_SEM2D.values()), 1)}
    n = big_ex * big_order + 1
    B = phc.Sem2dSolver(ex=big_ex, order=big_order, lx=20.0 * big_ex, dt=2e-4, f0=8.0, nsamples=big_steps + 1)
    chunk = B.autotune(steps_per_trial=5)
    B.set_timing(True)
    B.run(big_steps)
    kms = B.last_kernel_ms()
    fin = B.check_finite(B.i2) == 0
    B.close()
    ach = BYTES_PER_UPDATE * n * n / (kms * 1e-3) / 1e9
    out["large"] = {
        "workload": "SEM2D %d x %d elements of order %d (%d^2 nodes), %d steps" % (big_ex, big_ex, big_order, n,
                                                                                 big_steps),
        "kernel_ms_per_step": round(kms, 5), "value": round(n * n / (kms * 1e-3) / 1e6, 1),
        "unit": "Mnode-updates/s", "chunk": chunk, "finite": fin,
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": read_traffic(SEM2D_TRAFFIC_FILE)}}
    return out


def bench_pml(phc, n=512, steps=200):
    """FD with the reference's PML (-usePML, FDTDkernels.hpp:36-104) on the
    headline grid: width (int)(4 vmin/(2.5 f0)/dx) = 24 at vmin 1500 m/s,
    f0 10 Hz, dx 10 m; one marching launch per step (pml3D on the slab points,
    inner3D on the rest; HIP events around it).  Algorithmic bytes: 16 per
    inner point, 24 per slab point (pn, pnm1, vp^2dt^2, phi read; phi, pnew
    written)."""
    setup = phc.FdSetup(nx=n, model=("layered", 6, 20240512), f0=10.0, cfl=0.4, nsamples=steps + 1,
                        sponge_width=0)
    sol = phc.FdSolver(setup, arith=phc.ARITH_FP32)
    nd = sol.set_pml(f0=10.0)
    shape = sol.autotune(steps_per_trial=5)  # times the PML variant of the march
    sol.set_timing(True)
    sol.run(0, steps)
    kms = sol.last_kernel_ms()
    fin = sol.check_finite(sol.i2) == 0
    sol.close()
    inner = (n - 2 * nd) ** 3
    slab = n ** 3 - inner
    ach = (16 * inner + 24 * slab) / (kms * 1e-3) / 1e9
    return {"workload": "FD %d^3 radius 4 with a %d-point PML, %d steps, FP32" % (n, nd, steps),
            "pml_width": nd, "shape": list(shape), "ms_per_step": round(kms, 5), "value": round(n ** 3 / (kms * 1e-3) / 1e6, 1),
            "unit": "Mpoint-updates/s", "finite": fin,
            "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": read_traffic(PML_TRAFFIC_FILE)}}


FD_SIZES = (264, 336, 384, 401, 448)


def bench_fd_sizes(phc, sizes=FD_SIZES, steps=100):
    """The headline FD workload (radius 4, 20-point sponge, 6 seeded layers,
    Ricker 10 Hz, CFL 0.4, FP32) on other cubic grids: shape autotuned per
    size, then `steps` steps with HIP events around each launch.  16
    algorithmic bytes per point update, as for the headline."""
    out = {"workload": "FD n^3 radius 4, 20-pt sponge, 6 layers, %d steps, FP32" % steps, "sizes": {}}
    for n in sizes:
        setup = phc.FdSetup(nx=n, model=("layered", 6, 20240512), f0=10.0, cfl=0.4, nsamples=steps + 21)
        sol = phc.FdSolver(setup, arith=phc.ARITH_FP32)
        shape = sol.autotune(steps_per_trial=10)
        sol.set_timing(True)
        sol.run(0, 20)
        sol.run(20, steps)
        kms = sol.last_kernel_ms()
        sol.close()
        out["sizes"][str(n)] = {"shape": list(shape), "kernel_ms": round(kms, 5),
                                "value": round(n ** 3 / (kms * 1e-3) / 1e6, 1), "unit": "Mpoint-updates/s",
                                "roofline_frac": round(BYTES_PER_UPDATE * n ** 3 / (kms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)}
    return out


def bench_sem_elem(phc, steps=100, steps_p4=20):
    """The reference's CPU-runnable SEM case (50^3 elements of order 2, 101^3
    nodes, c = 1500, f0 = 5 Hz, dt = 1e-3) through the SEM step kernels:
    separable, element with atomic scatter, element with the 8-colour
    scatter, element write-once march (leapfrog fused), the 8-colour element
    kernel with its contractions on MFMA; then the same on
    100^3 elements of order 4 (401^3 nodes, dt = 2e-4, f0 = 8 Hz), with the
    roofline fraction at 16 B per node update."""
    out = {"workload": "SEM 50^3 elements of order 2 (101^3 nodes), %d steps, FP32" % steps}
    kernels = (("separable", 0), ("element_atomic", 1), ("element_colour", 2), ("element_march", 3),
               ("element_colour_mfma", 4))
    for name, kern in kernels:
        s = phc.SemSolver(ex=50, lx=2000.0, order=2, dt=1e-3, f0=5.0, nsamples=steps + 1, kernel=kern)
        if kern == 0:
            s.autotune(steps_per_trial=5)
        s.set_timing(True)
        s.run(steps)
        ms = s.last_kernel_ms()
        s.close()
        out[name] = {"ms_per_step": round(ms, 5), "value": round(101 ** 3 / (ms * 1e-3) / 1e6, 1),
                     "unit": "Mnode-updates/s"}
    p4 = {"workload": "SEM 100^3 elements of order 4 (401^3 nodes), %d steps, FP32" % steps_p4}
    for name, kern in kernels:
        s = phc.SemSolver(ex=100, lx=2000.0, order=4, dt=2e-4, f0=8.0, nsamples=steps_p4 + 6, kernel=kern)
        if kern == 0:
            s.autotune(steps_per_trial=5)
        s.set_timing(True)
        s.run(5)
        s.run(steps_p4)
        ms = s.last_kernel_ms()
        s.close()
        p4[name] = {"ms_per_step": round(ms, 5), "value": round(401 ** 3 / (ms * 1e-3) / 1e6, 1),
                    "unit": "Mnode-updates/s",
                    "roofline_frac": round(401 ** 3 * BYTES_PER_UPDATE / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)}
    out["order4"] = p4
    return out


# std of the Gaussian noise on the learned layer's training inputs: without it
# trained layers grow an instability in long rollouts in most runs
LB_TRAIN_NOISE = 0.02
# wide-sponge training shots (64^3 regions)
LB_TRAIN_SHOTS = 400


def bench_learned_step(phc, n=256, steps=300):
    """Time per step of the learned layer's grid (n + 8)^3 (correction pass,
    step, apply pass; an untrained layer costs the same) against the 40-point
    sponge's grid (n + 80)^3, graph-replayed, with the 28 receivers of the
    learned workload.  No training and no FP64 reference run."""
    import learned
    lb = phc.LearnedBoundary(seed=11)
    rec = learned.receivers_in_region(n)
    dt = learned.fd_dt()
    inter = np.array([15, 34, 78, 121, 215])
    out = {"workload": "FD %d^3 region, learned 4-point layer vs 40-point sponge, %d steps" % (n, steps)}
    pml_w = int(phc.lib().phc_fd_pml_width(1500.0, 10.0, learned.DX))
    for name, pad, sw, use_lb in (("learned", 4, 0, True), ("sponge40", 40, 40, False), ("pml", pml_w, 0, None)):
        N = n + 2 * pad
        vp = learned.layered_vp2dt2((N, N, N), pad, inter, np.linspace(1500.0, 4500.0, 6), dt)
        s = phc.FdSetup(N, N, N, L=4, dx=learned.DX, xs=n // 2 + pad, ys=n // 2 + pad, zs=n // 2 + pad,
                        f0=10.0, cfl=learned.CFL, sponge_width=sw, model=vp, nsamples=steps + 1)
        sol = phc.FdSolver(s, arith=phc.ARITH_FP32, init_value=0.0)
        if use_lb:
            sol.set_learned_boundary(lb, halo=0.0)
        elif use_lb is None:
            sol.set_pml(ndamp=pad, f0=10.0)
        shape = sol.autotune(steps_per_trial=5)
        sol.set_receivers(rec + pad)
        sol.run(0, steps)  # capture + first replay
        sol.synchronize()
        sol.i1, sol.i2 = 0, 1
        t0 = time.perf_counter()
        sol.run(0, steps)
        sol.synchronize()
        out[name] = {"grid": N, "shape": list(shape),
                     "ms_per_step": round((time.perf_counter() - t0) / steps * 1e3, 4)}
        sol.close()
    lb.close()
    out["speedup_learned_vs_sponge40"] = round(out["sponge40"]["ms_per_step"] / out["learned"]["ms_per_step"], 3)
    return out


def bench_learned(phc, n=256, steps=2000):
    """Learned thin boundary: trained from scratch on the GPU (400 wide-sponge
    shots, Adam), then one shot on an n^3 region: the 4-point learned layer,
    the 40-point sponge, the PML and the bare grid against a 2x-wide FP64
    domain, and an 8000-step run of the layer beside the bare grid."""
    import learned
    t0 = time.perf_counter()
    lb = phc.LearnedBoundary(seed=11)
    learned.training_data(lb, n_shots=LB_TRAIN_SHOTS, region=64)
    hist = learned.train(lb, steps=6000, noise=LB_TRAIN_NOISE)
    train_s = time.perf_counter() - t0
    r = learned.compare(lb, n=n, steps=steps)
    r.pop("truth_traces")
    r["rollout_8000"] = learned.rollout(lb, n=n, steps=8000)
    r["train_wall_s"] = round(train_s, 2)
    r["train_final_loss"] = float(hist[-1])
    r["train_input_noise"] = LB_TRAIN_NOISE
    r["samples_per_face"] = [int(c) for c in lb.counts()]
    lb.close()
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--arith", default="fp32", choices=["fp32", "ref"])
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched multi-shot measurement reported beside the headline")
    ap.add_argument("--workload", default="fd", choices=["fd", "batch", "sem", "sem2d", "pml", "semelem", "learned",
                                                        "learned-step", "fd-sizes"],
                    help="fd: the 512^3 headline; batch: only the batched multi-shot workload "
                         "(--steps steps, default 800); sem: only the SEM order-4 workload")
    ap.add_argument("--batch-shape", default=None,
                    help="tile_j,block_waves,chunk_i,blocks_per_cu for the batch (default: autotuned)")
    ap.add_argument("--no-baselines", action="store_true",
                    help="batch workload: skip the one-at-a-time runs")
    ap.add_argument("--no-fp64", action="store_true", help="sem workload: skip the FP64 run")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU decomposition "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import phc
    if phc.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.workload == "learned-step":
        r = bench_learned_step(phc)
        out = {"metric": METRIC, "value": r["speedup_learned_vs_sponge40"], "unit": "x (sponge step / learned step)",
               "n_gpus": 1, "steps": 300, "warmup": 0, "ms_per_step": r["learned"]["ms_per_step"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
               "data": "synthetic: seeded layered vp, untrained layer", "config": {"workload": r["workload"]},
               "learned_step": r}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "fd-sizes":
        r = bench_fd_sizes(phc)
        big = r["sizes"][str(FD_SIZES[-1])]
        out = {"metric": METRIC, "value": big["value"], "unit": big["unit"], "n_gpus": 1, "steps": 100,
               "warmup": 20, "ms_per_step": big["kernel_ms"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f32", "data": "synthetic: seeded 6-layer vp model, Ricker source",
               "config": {"workload": r["workload"]}, "fd_sizes": r}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "learned":
        r = bench_learned(phc)
        out = {"metric": METRIC, "value": r["learned"]["trace_rel_l2"], "unit": "receiver-trace rel. L2",
               "n_gpus": 1, "steps": r["steps"], "warmup": 0, "ms_per_step": r["learned"]["ms_per_step"],
               "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
               "data": "synthetic: seeded layered vp, Ricker sources",
               "config": {"workload": "learned 4-point layer vs 40-point sponge, 256^3 region, 2000 steps"},
               "learned_boundary": r}
        print(json.dumps(out), flush=True)
        return
    if args.workload in ("pml", "semelem"):
        r = bench_pml(phc) if args.workload == "pml" else bench_sem_elem(phc)
        cms = ctypes.c_double()  # plain 1 GiB copy: calibrates the rocprof byte counters
        phc.check(phc.lib().phc_copy_bandwidth(1 << 30, 5, ctypes.byref(cms)))
        r["copy_gbs"] = round(2 * (1 << 30) / (cms.value * 1e-3) / 1e9, 1)
        head = r if args.workload == "pml" else r["element_colour"]
        out = {"metric": METRIC, "value": head["value"], "unit": head["unit"], "n_gpus": 1, "steps": 0,
               "warmup": 0, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f32", "data": "synthetic",
               "config": {"workload": r["workload"]}, args.workload: r}
        if args.workload == "pml":
            out["roofline"] = r["roofline"]
        print(json.dumps(out), flush=True)
        return
    if args.workload == "sem2d":
        r = bench_sem2d(phc, big_steps=args.steps if "--steps" in sys.argv else 200)
        cms = ctypes.c_double()  # plain 1 GiB copy: calibrates the rocprof byte counters
        phc.check(phc.lib().phc_copy_bandwidth(1 << 30, 5, ctypes.byref(cms)))
        r["copy_gbs"] = round(2 * (1 << 30) / (cms.value * 1e-3) / 1e9, 1)
        big = r["large"]
        out = {"metric": METRIC, "value": big["value"], "unit": big["unit"], "n_gpus": 1,
               "steps": args.steps if "--steps" in sys.argv else 200, "warmup": 0,
               "ms_per_step": big["kernel_ms_per_step"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f32", "data": "synthetic: constant vp, Ricker source",
               "config": {"workload": big["workload"]}, "roofline": big["roofline"], "sem2d": r}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "sem":
        r = bench_sem_p4(phc, steps=args.steps if "--steps" in sys.argv else 2000, fp64=not args.no_fp64)
        cms = ctypes.c_double()  # plain 1 GiB copy: calibrates the rocprof byte counters
        phc.check(phc.lib().phc_copy_bandwidth(1 << 30, 5, ctypes.byref(cms)))
        r["copy_gbs"] = round(2 * (1 << 30) / (cms.value * 1e-3) / 1e9, 1)
        out = {"metric": METRIC, "value": r["value"], "unit": r["unit"], "n_gpus": 1, "steps": r["steps"],
               "warmup": 0, "ms_per_step": r["fp32_kernel_ms_per_step"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f32",
               "data": "synthetic: seeded Gaussian random vp per element, Ricker source",
               "config": {"workload": "SEM order 4, 100^3 elements of 20 m (401^3 nodes), 2000 steps"},
               "roofline": r["roofline"], "sem_p4": r}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "batch":
        shape = [int(v) for v in args.batch_shape.split(",")] if args.batch_shape else None
        nst = args.steps if "--steps" in sys.argv else 800
        r = bench_batch(phc, nsteps=nst, baselines=not args.no_baselines, shape=shape)
        out = {"metric": METRIC, "value": r["value"], "unit": "Mpoint-updates/s", "n_gpus": 1,
               "steps": nst, "warmup": nst, "ms_per_step": round(r["wall_s"] / nst * 1e3, 4),
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
               "data": "synthetic: seeded per-shot grids, 6-layer vp models and Ricker sources",
               "config": {"workload": "batched multi-shot FD: " + r["workload"]},
               "roofline": r["roofline"], "batch": r}
        print(json.dumps(out), flush=True)
        return
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
    sol.close()
    progress("FD 512^3 done: %.1f Mpt/s, kernel %.4f ms" % (value, kernel_ms))
    if not args.no_extras:
        out["batch"] = bench_batch(phc)
        progress("batched shots done")
        out["sem_p4"] = bench_sem_p4(phc)
        if not args.no_cpu:
            out["sem_p4"]["cpu_baseline"] = cpu_baseline_sem_p4()
        progress("SEM order 4 done")
        out["sem2d"] = bench_sem2d(phc)
        if not args.no_cpu:
            out["sem2d"]["cpu_baseline"] = cpu_baseline_sem2d()
        progress("SEM2D done")
        out["pml"] = bench_pml(phc)
        progress("PML done")
        out["fd_sizes"] = bench_fd_sizes(phc)
        progress("FD sizes done")
        out["sem_element_kernels"] = bench_sem_elem(phc)
        progress("SEM element kernels done")
        out["learned_boundary"] = bench_learned(phc)
        progress("learned boundary done")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(setup)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

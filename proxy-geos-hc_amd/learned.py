"""Learned thin absorbing boundary for the FD step: training data from
wide-sponge shots, GPU training, and the comparison on a region of interest.

No reference counterpart: the reference absorbs with a sponge
(src/fd/utilsFDTD/FDTDinit.hpp:319-396) or a PML (FDTDkernels.hpp:36-104).
The model and its kernels are in csrc/lb.hpp and csrc/mlp.hip; this module
drives them:

  training_data   wide-sponge FD shots (random source, 6-layer vp, random f0)
                  on a region plus margin plus a 40-point sponge; the layer
                  points just outside the region give inputs and exact targets
  train           Adam on the GPU (phc_lb_train), decaying learning rate
  compare         one shot on a region of n^3 points, run four ways:
                    learned  n + 2*4 points per side, the learned layer
                    sponge   n + 2*40 points per side, the 40-point sponge
                    none     n + 2*4 points per side, no absorbing boundary
                    truth    2n + 2*40 points per side in FP64
                  receiver-trace relative L2 against the truth, and wall time

Every grid uses h = 10 m, radius 4, dt from CFL 0.4 at vmax = 4500 m/s and
fields that start at 0 (the halo holds 0).
"""
import time

import numpy as np

import phc

DX = 10.0
VMIN, VMAX = 1500.0, 4500.0
CFL = 0.4
SPONGE = 40
LAYER = phc.LB_WIDTH
# the reference's PML width for this model and source (FDTDinit.hpp:167-172:
# 4 * vmin / (2.5 f0) / dx = 24 points at 1500 m/s, 10 Hz, 10 m)
PML_F0 = 10.0


def fd_dt(cfl=CFL, vmax=VMAX, L=4, dx=DX):
    c = phc.fd_coefs(L, dx)
    return np.float32(phc.lib().phc_fd_compute_dt(vmax, c, L, c, L, c, L, cfl))


def layered_vp2dt2(shape, k0, interfaces, velocities, dt):
    """vp^2 dt^2 (float32, IDX3) of horizontal layers along k: layer m spans
    region depths [interfaces[m-1], interfaces[m]); k0 = grid index of region
    depth 0 (depths outside the region continue the first / last layer)."""
    nx, ny, nz = shape
    depth = np.arange(nz) - k0
    layer = np.searchsorted(np.asarray(interfaces), depth, side="right")
    v = np.asarray(velocities, np.float32)[layer]
    col = (np.float32(v * v) * np.float32(dt)).astype(np.float32) * np.float32(dt)
    return np.broadcast_to(col.astype(np.float32)[None, None, :], shape).astype(np.float32).reshape(-1)


def random_layers(rng, region, nlayers=6):
    inter = np.sort(rng.choice(np.arange(4, region - 4), nlayers - 1, replace=False))
    vel = np.sort(rng.uniform(VMIN, VMAX, nlayers))
    vel[0], vel[-1] = VMIN + rng.uniform(0, 300), VMAX - rng.uniform(0, 300)
    return inter, np.sort(vel)


def training_data(lb, n_shots=200, region=64, margin=8, steps=600, collect_every=6, keep=0.006, s_min=1e-6,
                  f0_range=(6.0, 14.0), seed=2024, capacity=4_000_000, log=None):
    """Fill lb's sample store from n_shots wide-sponge shots."""
    rng = np.random.default_rng(seed)
    dt = fd_dt()
    N = region + 2 * (margin + SPONGE)
    o = margin + SPONGE
    lb.set_capacity(capacity)
    t0 = time.time()
    for shot in range(n_shots):
        inter, vel = random_layers(rng, region)
        src = [int(o + v) for v in rng.integers(8, region - 8, 3)]
        f0 = float(rng.uniform(*f0_range))
        vp = layered_vp2dt2((N, N, N), o, inter, vel, dt)
        s = phc.FdSetup(N, N, N, L=4, dx=DX, xs=src[0], ys=src[1], zs=src[2], f0=f0, cfl=CFL,
                        sponge_width=SPONGE, model=vp, nsamples=steps + 1)
        sol = phc.FdSolver(s, arith=phc.ARITH_FP32, init_value=0.0)
        for it in range(steps):
            sol.step(it)  # eager launches: no graph to capture for short runs
            if (it + 1) % collect_every == 0:
                lb.collect(sol, sol.i2, (o, o, o), (region, region, region), halo=0.0, keep=keep, s_min=s_min,
                           seed=seed * 100003 + shot * 1009 + it)
        sol.close()
        if log and (shot + 1) % 20 == 0:
            log("  shot %d/%d  samples per face %s  %.1fs" % (shot + 1, n_shots, list(lb.counts()),
                                                                time.time() - t0))
    return lb.counts()


def onpolicy_data(lb, n_shots=60, region=64, margin=8, steps=600, collect_every=6, keep=0.006, s_min=1e-6,
                  f0_range=(6.0, 14.0), seed=4048, log=None):
    """Append on-policy samples: each shot runs on the region itself with the
    model's current weights attached and, in lockstep, on the wide grid; the
    thin run's own layer states are labelled with the wide run's field beyond
    the region (phc_lb_collect_onpolicy)."""
    rng = np.random.default_rng(seed)
    dt = fd_dt()
    N = region + 2 * (margin + SPONGE)
    o = margin + SPONGE
    t0 = time.time()
    for shot in range(n_shots):
        inter, vel = random_layers(rng, region)
        src = [int(v) for v in rng.integers(8, region - 8, 3)]
        f0 = float(rng.uniform(*f0_range))
        ws = phc.FdSetup(N, N, N, L=4, dx=DX, xs=src[0] + o, ys=src[1] + o, zs=src[2] + o, f0=f0, cfl=CFL,
                         sponge_width=SPONGE, model=layered_vp2dt2((N, N, N), o, inter, vel, dt),
                         nsamples=steps + 1)
        ts = phc.FdSetup(region, region, region, L=4, dx=DX, xs=src[0], ys=src[1], zs=src[2], f0=f0, cfl=CFL,
                         sponge_width=0, model=layered_vp2dt2((region,) * 3, 0, inter, vel, dt),
                         nsamples=steps + 1)
        wide = phc.FdSolver(ws, arith=phc.ARITH_FP32, init_value=0.0)
        thin = phc.FdSolver(ts, arith=phc.ARITH_FP32, init_value=0.0)
        thin.set_learned_boundary(lb, halo=0.0)
        for it in range(steps):
            wide.step(it)
            thin.step(it)
            if (it + 1) % collect_every == 0:
                lb.collect_onpolicy(thin, wide, (o, o, o), halo=0.0, keep=keep, s_min=s_min,
                                    seed=seed * 100003 + shot * 1009 + it)
        wide.close()
        thin.close()
        if log and (shot + 1) % 20 == 0:
            log("  on-policy shot %d/%d  samples per face %s  %.1fs" % (shot + 1, n_shots, list(lb.counts()),
                                                                          time.time() - t0))
    return lb.counts()


def train(lb, steps=6000, batch_per_face=4096, lr=2e-3, lr_final=1e-4, seed=7, log=None, noise=0.0):
    """Adam with a geometric learning-rate decay in 6 stages; returns the loss
    history.  noise: std of the Gaussian noise on the training inputs."""
    lb.set_train_noise(noise)
    hist = []
    stages = 6
    for k in range(stages):
        rate = lr * (lr_final / lr) ** (k / (stages - 1))
        h = lb.train(steps // stages, batch_per_face=batch_per_face, lr=rate, seed=seed * 31 + k)
        hist.append(h)
        if log:
            log("  stage %d lr %.2e  loss %.3e -> %.3e" % (k, rate, h[0], h[-1]))
    return np.concatenate(hist)


def receivers_in_region(n, inset=8, per_side=3):
    """Receivers on the six faces of the region (inset points inside), per_side^2
    on each, plus the centre line; region coordinates."""
    pos = np.linspace(inset, n - 1 - inset, per_side).astype(int)
    rec = []
    for a in range(3):
        for side in (inset, n - 1 - inset):
            for u in pos:
                for v in pos:
                    c = [0, 0, 0]
                    c[a] = side
                    c[(a + 1) % 3] = u
                    c[(a + 2) % 3] = v
                    rec.append(c)
    for d in (n // 4, n // 2 + n // 8):
        rec.append([d, n // 2, n // 2])
    return np.unique(np.array(rec, np.int32), axis=0)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def compare(lb, n=256, steps=2000, f0=10.0, seed=20240512, src=None, with_none=True, log=None,
            layer_damping=0.0, truth=None):
    """One shot on an n^3 region: learned layer vs 40-point sponge (vs no
    absorbing boundary) against a 2n-wide FP64 domain; returns a dict."""
    rng = np.random.default_rng(seed)
    inter, vel = np.sort(rng.choice(np.arange(8, n - 8), 5, replace=False)), np.linspace(VMIN, VMAX, 6)
    dt = fd_dt()
    src = src or (n // 2, n // 2, n // 2)
    rec = receivers_in_region(n)
    out = {"region": n, "steps": steps, "f0": f0, "dt": float(dt), "receivers": int(len(rec)),
           "layers_k": [int(v) for v in inter], "velocities": [float(v) for v in vel]}

    def setup(pad, sw):
        N = n + 2 * pad
        vp = layered_vp2dt2((N, N, N), pad, inter, vel, dt)
        return phc.FdSetup(N, N, N, L=4, dx=DX, xs=src[0] + pad, ys=src[1] + pad, zs=src[2] + pad, f0=f0,
                           cfl=CFL, sponge_width=sw, model=vp, nsamples=steps + 1)

    def run32(pad, sw, model):
        s = setup(pad, sw)
        sol = phc.FdSolver(s, arith=phc.ARITH_FP32, init_value=0.0)
        if model == "pml":
            sol.set_pml(ndamp=pad, f0=PML_F0)
        elif model is not None:
            sol.set_learned_boundary(model, halo=0.0, layer_damping=layer_damping)
        sol.autotune(steps_per_trial=5)
        sol.set_receivers(rec + pad)
        sol.run(0, steps)               # warm-up: captures the run's graph
        z = np.zeros_like(sol.field(0))
        sol.set_field(0, z)
        sol.set_field(1, z)
        sol.i1, sol.i2 = 0, 1           # the same start: the timed run replays the graph
        sol.synchronize()
        t0 = time.perf_counter()
        _, tr = sol.run(0, steps)
        sol.synchronize()
        wall = time.perf_counter() - t0
        fin = sol.check_finite(sol.i2) == 0
        pts = s.nx * s.ny * s.nz
        sol.close()
        return tr, wall, fin, pts

    # truth: 2n region in FP64 with its own sponge, region centred
    pad_t = n // 2 + SPONGE
    if truth is None:
        t0 = time.perf_counter()
        s64 = setup(pad_t, SPONGE)
        f64 = phc.Fd64Solver(s64, init_value=0.0)
        f64.set_receivers(rec + pad_t)
        _, tr_true = f64.run(steps)
        f64.close()
        out["truth_wall_s"] = round(time.perf_counter() - t0, 3)
        if log:
            log("  truth (FP64, %d^3): %.1fs" % (n + 2 * pad_t, out["truth_wall_s"]))
    else:
        tr_true = truth
    out["truth_traces"] = tr_true
    out["layer_damping"] = layer_damping
    pml_w = int(phc.lib().phc_fd_pml_width(VMIN, PML_F0, DX))
    runs = [("learned", LAYER, 0, lb), ("sponge40", SPONGE, SPONGE, None), ("pml", pml_w, 0, "pml")]
    if with_none:
        runs.append(("none", LAYER, 0, None))
    # receiver groups: the centre of each face's 3x3 grid, and the interior line
    mid = np.linspace(8, n - 9, 3).astype(int)[1]
    on_face = [(r == 8) | (r == n - 9) for r in rec.T]
    n_face = np.sum(on_face, axis=0)
    centre = (n_face == 1) & np.all((rec == mid) | np.array(on_face).T, axis=1)
    interior = n_face == 0
    out["receiver_groups"] = {"face_centres": int(centre.sum()), "interior": int(interior.sum())}
    for name, pad, sw, model in runs:
        tr, wall, fin, pts = run32(pad, sw, model)
        out[name] = {"grid": n + 2 * pad, "points": pts, "wall_s": round(wall, 4),
                     "ms_per_step": round(wall / steps * 1e3, 4), "finite": bool(fin),
                     "trace_rel_l2": rel_l2(tr, tr_true) if fin else None,
                     "trace_rel_l2_face_centres": rel_l2(tr[centre], tr_true[centre]) if fin else None,
                     "trace_rel_l2_interior": rel_l2(tr[interior], tr_true[interior]) if fin else None}
        if log:
            log("  %-9s grid %d^3: %.3f s, trace rel L2 %s (face centres %s, interior %s)" % (
                name, n + 2 * pad, wall, out[name]["trace_rel_l2"], out[name]["trace_rel_l2_face_centres"],
                out[name]["trace_rel_l2_interior"]))
    out["speedup_learned_vs_sponge40"] = round(out["sponge40"]["wall_s"] / out["learned"]["wall_s"], 3)
    out["speedup_learned_vs_pml"] = round(out["pml"]["wall_s"] / out["learned"]["wall_s"], 3)
    return out


def rollout(lb, n=256, steps=8000, every=1000, f0=10.0, seed=20240512, layer_damping=0.0):
    """One shot on an n^3 region run long: the largest |p| on the grid after
    every `every` steps, with the learned layer and on the bare grid (its edge
    reflects, the energy stays).  The Ricker source has faded after a few
    hundred steps, so with an absorbing layer the maximum keeps falling, while
    a layer that feeds energy back makes it grow past the bare grid's."""
    rng = np.random.default_rng(seed)
    inter, vel = np.sort(rng.choice(np.arange(8, n - 8), 5, replace=False)), np.linspace(VMIN, VMAX, 6)
    dt = fd_dt()
    out = {"region": n, "steps": steps, "every": every}
    for name, model in (("learned", lb), ("none", None)):
        N = n + 2 * LAYER
        vp = layered_vp2dt2((N, N, N), LAYER, inter, vel, dt)
        s = phc.FdSetup(N, N, N, L=4, dx=DX, xs=n // 2 + LAYER, ys=n // 2 + LAYER, zs=n // 2 + LAYER, f0=f0,
                        cfl=CFL, sponge_width=0, model=vp, nsamples=steps + 1)
        sol = phc.FdSolver(s, arith=phc.ARITH_FP32, init_value=0.0)
        if model is not None:
            sol.set_learned_boundary(model, halo=0.0, layer_damping=layer_damping)
        maxima = []
        for it in range(0, steps, every):
            sol.run(it, min(every, steps - it))
            f = sol.current()
            maxima.append(float(np.abs(f).max()) if np.isfinite(f).all() else float("inf"))
        out[name] = maxima
        sol.close()
    out["bounded"] = bool(max(out["learned"]) <= 2.0 * max(out["none"]))
    return out


def save_params(lb, path):
    lb.params().astype("<f4").tofile(path)


def load_params(lb, path):
    th = np.fromfile(path, dtype="<f4")
    if th.size != phc.LB_FACES * phc.LB_NP:
        raise ValueError("%s holds %d values, expected %d" % (path, th.size, phc.LB_FACES * phc.LB_NP))
    lb.set_params(th.reshape(phc.LB_FACES, phc.LB_NP))


def main():
    import argparse
    import json
    import os
    ap = argparse.ArgumentParser(description="train and evaluate the learned thin boundary")
    ap.add_argument("--shots", type=int, default=200)
    ap.add_argument("--region", type=int, default=64)
    ap.add_argument("--train-steps", type=int, default=6000)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "lb_weights_r4.f32"))
    ap.add_argument("--eval-n", type=int, default=256)
    ap.add_argument("--eval-steps", type=int, default=2000)
    ap.add_argument("--skip-train", action="store_true")
    ap.add_argument("--rounds", type=int, default=2, help="on-policy rounds after the first training")
    ap.add_argument("--onpolicy-shots", type=int, default=60)
    args = ap.parse_args()
    log = lambda m: print(m, flush=True)
    lb = phc.LearnedBoundary(seed=11)
    if args.skip_train:
        load_params(lb, args.out)
    else:
        log("training data: %d shots" % args.shots)
        training_data(lb, n_shots=args.shots, region=args.region, log=log)
        log("training")
        hist = train(lb, steps=args.train_steps, log=log)
        for r in range(args.rounds):
            log("on-policy round %d: %d shots" % (r + 1, args.onpolicy_shots))
            onpolicy_data(lb, n_shots=args.onpolicy_shots, region=args.region, seed=4048 + 17 * r, log=log)
            hist = train(lb, steps=args.train_steps // 2, lr=5e-4, lr_final=5e-5, seed=100 + r, log=log)
        save_params(lb, args.out)
        log("saved %s (final minibatch loss %.3e)" % (args.out, hist[-1]))
    log("comparison on %d^3, %d steps" % (args.eval_n, args.eval_steps))
    res = compare(lb, n=args.eval_n, steps=args.eval_steps, log=log)
    res.pop("truth_traces")
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

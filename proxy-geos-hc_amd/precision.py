"""SEM order-4 precision study: the FP32 time loop against the FP64 one on the
same inputs (FP64 setup in both: GLL basis, lumped mass, dt^2 W/M).

Workload: 100 x 100 x 100 hex elements of 20 m at GLL order 4 (401^3 nodes),
per-element wave speed from a seeded smooth Gaussian random field scaled to
2000-4000 m/s, Ricker source (f0 = 8 Hz) at the centre, dt = 2e-4 s (Courant
number 0.23 on the 3.45 m minimum GLL spacing), 2000 steps, exact GLL weights.
Reported: relative L2 of the final field, of the source trace and of the
receiver traces (8 nodes on a line at depth 100 m through the centre), and the
FP32 kernel time per step (separable kernel, HIP events).

No reference counterpart: the reference runs in FP32 only (SEMsolver.cpp:21-125).
"""
import time

import numpy as np

import phc


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def sem_p4_study(ex=100, order=4, h=20.0, steps=2000, dt=2e-4, f0=8.0, seed=20240512, kernel=phc.SEM_SEPARABLE,
                 log=None, fp64=True):
    lx = h * ex
    model = phc.sem_model_gaussian(ex, ex, ex, sigma=5.0, vmin=2000.0, vmax=4000.0, seed=seed)
    centre = (lx / 2 + 0.5 * h / order, lx / 2 + 0.5 * h / order, lx / 2 + 0.5 * h / order)
    kw = dict(ex=ex, lx=lx, order=order, dt=dt, f0=f0, model=model, nsamples=steps, src_point=centre)
    nx = ex * order + 1
    iz = int(round(100.0 / (h / order)))
    iy = nx // 2
    rec = np.array([ix + iz * nx + iy * nx * nx for ix in np.linspace(nx // 8, nx - 1 - nx // 8, 8).astype(int)],
                   np.int32)
    out = {"elements": ex ** 3, "order": order, "nodes": nx ** 3, "steps": steps, "dt": dt, "f0": f0,
           "h": h, "receivers": len(rec)}
    t0 = time.perf_counter()
    s32 = phc.SemSolver(weights_ref=False, kernel=kernel, **kw)
    s32.autotune(steps_per_trial=5)
    s32.set_receivers(rec)
    s32.set_timing(True)
    tr32, rv32 = s32.run(steps)
    out["fp32_kernel_ms_per_step"] = round(s32.last_kernel_ms(), 5)
    f32 = s32.current().astype(np.float64)
    s32.close()
    out["fp32_wall_s"] = round(time.perf_counter() - t0, 2)
    if log:
        log("  FP32 done (%.1fs, %.4f ms/step)" % (out["fp32_wall_s"], out["fp32_kernel_ms_per_step"]))
    if not fp64:
        return out
    t0 = time.perf_counter()
    s64 = phc.Sem64Solver(weights_ref=False, **kw)
    s64.set_receivers(rec)
    tr64, rv64 = s64.run(steps)
    f64 = s64.field(True)
    s64.close()
    out["fp64_wall_s"] = round(time.perf_counter() - t0, 2)
    if log:
        log("  FP64 done (%.1fs)" % out["fp64_wall_s"])
    out["field_rel_l2"] = rel_l2(f32, f64)
    out["source_trace_rel_l2"] = rel_l2(tr32, tr64)
    out["receiver_traces_rel_l2"] = rel_l2(rv32, rv64)
    out["field_max_abs"] = float(np.max(np.abs(f64)))
    out["finite"] = bool(np.all(np.isfinite(f32)))
    return out


def main():
    import argparse
    import json
    ap = argparse.ArgumentParser(description="SEM order-4 FP32 vs FP64 precision study")
    ap.add_argument("--ex", type=int, default=100)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--kernel", default="separable", choices=["separable", "element"])
    a = ap.parse_args()
    k = phc.SEM_SEPARABLE if a.kernel == "separable" else phc.SEM_ELEMENT
    res = sem_p4_study(ex=a.ex, steps=a.steps, kernel=k, log=lambda m: print(m, flush=True))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

"""FP32 drift at the benchmark's length: the all-FP32 step (the one bench.py
times) against the reference-rounding step, which is bit-exact with the
oracle, over the full 1000 steps of the 512^3 layered workload (6 seeded
layers 1500-4500 m/s, f0 = 10 Hz, dt from CFL 0.4, 20-point sponge), and the
same for the PML step (the reference's default PML width at this model).
In 1000 steps the wave travels at most ~135 points, short of the absorbing
layers of the 512^3 grid, so the same runs on a 192^3 grid (layers reached
after ~300 steps) cover the sponge and PML arithmetic.

Bar: relative L2 <= 1e-4 (the stated parity tolerance) on the final field, the
source-point trace and 8 receiver traces.  The measured drift is printed."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NSTEPS = 1000


def receivers(n):
    # 8 receivers on a line 40 points above the source (400-690 m away)
    return np.array([[n // 2 - 56 + 16 * r, n // 2, n // 2 - 40] for r in range(8)], np.int32)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_both(gpu, s, pml):
    out = {}
    for name, arith in (("ref", gpu.ARITH_REF), ("fp32", gpu.ARITH_FP32)):
        print("[fd-drift] %s %s %d^3 x %d" % ("pml" if pml else "sponge", name, s.nx, NSTEPS), flush=True)
        sol = gpu.FdSolver(s, arith=arith)
        if pml:
            sol.set_pml(f0=s.f0)
        sol.set_receivers(receivers(s.nx))
        src, rec = sol.run(0, NSTEPS)
        out[name] = (src.copy(), rec.copy(), sol.current())
        sol.close()
    return out


@pytest.mark.parametrize("n,pml", [(512, False), (512, True), (192, False), (192, True)])
def test_fp32_drift_over_benchmark_length(gpu, n, pml):
    s = gpu.FdSetup(nx=n, model=("layered", 6, 20240512), f0=10.0, cfl=0.4, nsamples=NSTEPS,
                    sponge_width=0 if pml else 20)
    r = run_both(gpu, s, pml)
    e_field = rel_l2(r["fp32"][2], r["ref"][2])
    e_src = rel_l2(r["fp32"][0], r["ref"][0])
    e_rec = rel_l2(r["fp32"][1], r["ref"][1])
    print("[fd-drift] %s %d^3: field %.3e source %.3e receivers %.3e"
          % ("pml" if pml else "sponge", n, e_field, e_src, e_rec), flush=True)
    assert np.isfinite(r["fp32"][2]).all()
    assert np.abs(r["ref"][1]).max() > 0  # the wave reached the receivers
    assert e_field <= 1e-4
    assert e_src <= 1e-4
    assert e_rec <= 1e-4

"""The reference's CPU-runnable SEM case in full on the GPU: a cube of 50^3
hex elements of 40 m, GLL order 2 (101^3 nodes), c = 1500 m/s, Ricker 5 Hz
at the centre (the element holding (1001, 1001, 1001), SEMproxy.cpp:96-106),
dt = 1e-3 s, 1500 steps, 8 receivers at x = 200 (r+1) m, y = 1000 m,
z = 100 m.  Every FP32 kernel against the FP64 loop over all 1500 steps, and
against the FP32 reference-semantics oracle over a prefix the oracle finishes
in seconds."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KW = dict(ex=50, lx=2000.0, order=2, dt=1e-3, f0=5.0, nsamples=1500, src_point=(1001.0, 1001.0, 1001.0))


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def receivers(nx=101, nz=101):
    # node (ix, iy, iz) = ix + iz nx + iy nx nz; 20 m node spacing
    return np.array([(10 * (r + 1)) + 5 * nx + 50 * nx * nz for r in range(8)], np.int32)


@pytest.fixture(scope="module")
def fp64(gpu):
    print("\n[config0] FP64 run, 1500 steps", flush=True)
    s = gpu.Sem64Solver(weights_ref=True, **KW)
    s.set_receivers(receivers())
    src, rec = s.run(1500)
    return src, rec, s.field(True)


@pytest.mark.parametrize("kernel", [0, 1, 2])
def test_full_run_against_fp64(gpu, fp64, kernel):
    print("\n[config0] FP32 kernel %d, 1500 steps" % kernel, flush=True)
    s = gpu.SemSolver(kernel=kernel, weights_ref=True, **KW)
    assert s.src_element == 25 + 25 * 50 + 25 * 2500
    s.set_receivers(receivers())
    src, rec = s.run(1500)
    src64, rec64, f64 = fp64
    assert np.abs(rec64).max() > 0
    # the target bar is 1e-4 relative L2 on traces and field
    assert rel(src, src64) < 1e-5
    assert rel(rec, rec64) < 1e-4
    assert rel(s.current(), f64) < 1e-4


def test_prefix_against_reference_semantics(gpu, oracle):
    print("\n[config0] 40-step prefix vs the oracle", flush=True)
    O = oracle.SemProblem(**KW)
    otr, orec = O.run(40, receivers=receivers())
    s = gpu.SemSolver(kernel=0, weights_ref=True, **KW)
    s.set_receivers(receivers())
    src, rec = s.run(40)
    assert rel(src, otr) < 1e-5
    assert rel(s.current(), O.current()) < 1e-5

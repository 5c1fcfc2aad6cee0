"""FP64 SEM loop on the GPU (phc_sem64_*) against the CPU FP64 restatement
(oracle/sem_oracle.c orc_sem64_*, element-by-element sum factorisation with
exact box geometry), and the FP32 separable path against it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("order,ex,steps", [(2, 6, 60), (4, 5, 60), (3, 4, 40), (1, 7, 50)])
def test_fp64_gpu_matches_fp64_oracle(gpu, oracle, order, ex, steps):
    print("\n[sem64] order %d ex %d" % (order, ex), flush=True)
    rng = np.random.default_rng(order)
    model = rng.uniform(1500, 3000, ex ** 3).astype(np.float32)
    kw = dict(ex=ex, lx=200.0 * ex, order=order, dt=2e-4, f0=25.0, model=model, nsamples=steps,
              src_point=(100.0 * ex + 1, 100.0 * ex + 1, 100.0 * ex + 1))
    P = oracle.SemProblem(fp64=True, **kw)
    g = gpu.Sem64Solver(weights_ref=False, **kw)
    assert np.array_equal(P.rhs, g.rhs) and P.src_element == g.src_element
    nodes = np.array([P.src_node, P.src_node + 1, P.nN // 2 + 3, P.nN - P.mesh.nx - 2], np.int32)
    g.set_receivers(nodes)
    tr, rv = g.run(steps)
    otr, orv = P.run(steps, receivers=nodes)
    assert rel(tr, otr) < 1e-12
    assert rel(rv, orv) < 1e-12
    assert rel(g.field(True), P.current()) < 1e-12
    assert rel(g.field(False), P.previous()) < 1e-12


def test_fp32_separable_tracks_fp64(gpu):
    """The FP32 separable path against the FP64 path on the same inputs:
    the FP32 rounding of a few hundred steps stays near single precision."""
    rng = np.random.default_rng(9)
    ex, order, steps = 8, 4, 300
    model = rng.uniform(2000, 4000, ex ** 3).astype(np.float32)
    kw = dict(ex=ex, lx=20.0 * ex, order=order, dt=2e-4 / 8, f0=200.0, model=model, nsamples=steps,
              src_point=(10.0 * ex + 0.5, 10.0 * ex + 0.5, 10.0 * ex + 0.5))
    s32 = gpu.SemSolver(weights_ref=False, **kw)
    s64 = gpu.Sem64Solver(weights_ref=False, **kw)
    t32, _ = s32.run(steps)
    t64, _ = s64.run(steps)
    assert np.max(np.abs(t64)) > 0
    assert rel(t32, t64) < 1e-4
    assert rel(s32.current(), s64.field(True)) < 1e-4

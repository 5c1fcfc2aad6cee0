"""SEM2D step on the GPU (phc_sem2d_*, csrc/sem2d.hip) against the SEM2D
oracle (reference semantics of SEMsolver.cpp:21-125 with SEM2D) and the
reference's published 2D runs."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


# Both FP32 paths carry about 1e-5 of rounding after a few hundred steps
# near a strong point source: the same separable scheme evaluated in FP64 lies
# 8e-6 (order 1) to 3e-6 (order 3) from the FP32 oracle on the 11 x 9 case
# below.  Bar: 3e-5 relative L2 (the target bar is 1e-4).
TOL = 3e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def layered_model(ex, ez, seed):
    rng = np.random.default_rng(seed)
    rows = rng.uniform(1500.0, 3500.0, ez).astype(np.float32)
    return np.repeat(rows, ex).astype(np.float32)


def pair(gpu, oracle, ex, ez, p, steps, src_element, seed=3, dt=5e-4, chunk=0, h=40.0):
    model = layered_model(ex, ez, seed)
    kw = dict(ex=ex, ez=ez, order=p, lx=h * ex, lz=h * ez, dt=dt, f0=12.0, model=model,
              nsamples=steps + 1, src_element=src_element)
    G = gpu.Sem2dSolver(chunk=chunk, **kw)
    O = oracle.Sem2dProblem(**kw)
    assert np.array_equal(G.rhs, O.rhs)
    return G, O


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_fused_run_matches_oracle(gpu, oracle, p):
    # source in an edge element: the wave reaches the paraxial edges early
    ex, ez, steps = 11, 9, 240
    G, O = pair(gpu, oracle, ex, ez, p, steps, src_element=1 + 2 * ex)
    nodes = [G.src_node, 0, G.nx - 1, G.nn - 1, G.nx * (G.nz // 2), G.nx * (G.nz // 2) + G.nx // 2]
    G.set_receivers(nodes)
    src, rec = G.run(steps)
    osrc, orec = O.run(steps, receivers=nodes)
    assert rel(src, osrc) < TOL
    assert rel(rec, orec) < TOL
    assert rel(G.current(), O.current()) < TOL
    assert rel(G.previous(), O.previous()) < TOL
    assert G.check_finite(G.i2) == 0


def test_step_by_step_matches_oracle(gpu, oracle):
    G, O = pair(gpu, oracle, 8, 13, 3, 60, src_element=5 + 6 * 8)
    for _ in range(60):
        G.step()
    O.run(60)
    assert rel(G.current(), O.current()) < TOL
    assert rel(G.previous(), O.previous()) < TOL


@pytest.mark.parametrize("chunk", [1, 3, 64, 1 << 20])
def test_rows_per_wave_do_not_change_results(gpu, oracle, chunk):
    # every node's arithmetic is the same whatever the march length
    ref = gpu.Sem2dSolver(ex=70, ez=9, order=4, lx=2800.0, lz=360.0, dt=5e-4, nsamples=80,
                          src_element=3 + 4 * 70, chunk=2)
    ref.run(79)
    G = gpu.Sem2dSolver(ex=70, ez=9, order=4, lx=2800.0, lz=360.0, dt=5e-4, nsamples=80,
                        src_element=3 + 4 * 70, chunk=chunk)
    G.run(79)
    assert np.array_equal(G.current(), ref.current())
    assert np.array_equal(G.previous(), ref.previous())


def test_64bit_offsets_give_the_same_bits(gpu):
    kw = dict(ex=21, ez=17, order=4, lx=840.0, lz=680.0, dt=5e-4, nsamples=60, src_element=3 + 5 * 21)
    A = gpu.Sem2dSolver(**kw)
    B = gpu.Sem2dSolver(index64=True, **kw)
    a_src, _ = A.run(50)
    b_src, _ = B.run(50)
    assert np.array_equal(a_src, b_src)
    assert np.array_equal(A.current(), B.current()) and np.array_equal(A.previous(), B.previous())


def test_mass_and_boundary_term_match_oracle(gpu, oracle):
    ex, ez, p = 9, 7, 4
    G, O = pair(gpu, oracle, ex, ez, p, 1, src_element=0)
    O.run(1)
    Mo, _, sho = O.mass_y_sh()
    Mg, shg = G.mass()
    # the reference forms the Jacobian from FP32 node coordinates (sums of
    # x l'(xi) that cancel down to h/2), about 1e-6 relative; the device
    # setup is FP64 throughout
    assert np.allclose(Mg, Mo, rtol=1e-5, atol=0)
    want = np.zeros(G.nn, np.float32)
    want[O.bnodes] = sho
    assert np.allclose(shg, want, rtol=1e-5, atol=0)
    assert np.all(shg[O.interior] == 0)


def test_graph_replays_match_eager(gpu):
    kw = dict(ex=20, ez=16, order=3, lx=800.0, lz=640.0, dt=5e-4, nsamples=200, src_element=7 + 8 * 20)
    A = gpu.Sem2dSolver(**kw)
    B = gpu.Sem2dSolver(**kw)
    for _ in range(3):  # three captured runs of 50 steps, eager steps alongside
        A.run(50)
        for _ in range(50):
            B.step()
        A.synchronize()
        assert np.array_equal(A.current(), B.current())
        assert np.array_equal(A.previous(), B.previous())


def test_published_2d_run(gpu, oracle):
    # the reference's published SEM configuration (test_20240512 logs):
    # 65 x 65 elements of order 3, c = 1500, dt = 1e-3, f0 = 10 Hz, source at
    # element 2178; prints after steps 0, 100, ..., 900
    G = gpu.Sem2dSolver(ex=65, order=3, lx=2000.0, dt=1e-3, f0=10.0, nsamples=1000, src_element=2178)
    src, _ = G.run(901)
    O = oracle.Sem2dProblem(ex=65, order=3, lx=2000.0, dt=1e-3, f0=10.0, nsamples=1000, src_element=2178)
    osrc, _ = O.run(901)
    assert rel(src, osrc) < TOL
    # after 900 steps the wave has left through the paraxial edges; against
    # the same step in FP64 the reference-semantics FP32 loop is 7.6e-5 off
    # (relative L2 of the remaining field) and the device loop, which sums
    # its stencil over differences from the centre, several times closer
    O64 = oracle.Sem2d64Problem(ex=65, order=3, lx=2000.0, dt=1e-3, f0=10.0, nsamples=1000, src_element=2178)
    t64 = O64.run(901)
    e_dev, e_ref = rel(G.current(), O64.current()), rel(O.current(), O64.current())
    print("[sem2d] field vs FP64: device %.2e, reference semantics %.2e" % (e_dev, e_ref), flush=True)
    assert e_dev < 2e-5 and e_dev < e_ref / 3
    assert rel(src, t64) < 1e-5
    assert rel(G.current(), O.current()) < 1e-4
    d = json.load(open(os.path.join(GOLDEN, "published_traces.json")))
    vals = np.array([[float(v) for v in r["values"]] for k, r in d.items()
                     if "_sem_" in k and "_OMP_" not in k])
    mine = src[::100]
    # the published prints share the reference's FP32 rounding (the FP64 run
    # prints 169.032 at step 100 and -0.33427 at step 900 against 169.033 and
    # -0.33441..-0.33447 published): each print here is within 1e-5 of the
    # trace's peak of the nearest published value
    peak = np.abs(src).max()
    for k in range(10):
        assert np.min(np.abs(vals[:, k] - mine[k])) <= 1e-5 * peak, (k, mine[k], vals[:, k].min(), vals[:, k].max())


def test_autotune_keeps_state(gpu):
    G = gpu.Sem2dSolver(ex=30, ez=30, order=3, lx=1200.0, dt=5e-4, nsamples=40, src_element=15 + 15 * 30)
    G.run(20)
    before = (G.current().copy(), G.previous().copy())
    c = G.autotune(steps_per_trial=2)
    assert c >= 1
    assert np.array_equal(G.current(), before[0]) and np.array_equal(G.previous(), before[1])
    R = gpu.Sem2dSolver(ex=30, ez=30, order=3, lx=1200.0, dt=5e-4, nsamples=40, src_element=15 + 15 * 30)
    R.run(20)
    G.run(19)
    R.run(19)
    assert np.array_equal(G.current(), R.current())


def test_rejects_bad_parameters(gpu):
    with pytest.raises(RuntimeError, match="ey = 0"):
        p = gpu.SemParams()
        p.ex, p.ey, p.ez, p.order = 4, 4, 4, 2
        import ctypes
        h = ctypes.c_void_p()
        gpu.check(gpu.lib().phc_sem2d_create(ctypes.byref(p), ctypes.byref(h)))
    with pytest.raises(RuntimeError, match="order"):
        gpu.Sem2dSolver(ex=4, order=6, nsamples=2, src_element=0)


def test_full_size_8001_rows_per_wave_and_offsets(gpu):
    # the benchmark's large mesh (2000 x 2000 elements of order 4, 8001^2
    # nodes), source in the element beside a corner so that the pulse reaches
    # both paraxial edges near it: marches of 2, 32 and all rows per
    # wavefront and the 64-bit-offset kernel give the same bits
    kw = dict(ex=2000, order=4, lx=40000.0, dt=2e-4, f0=25.0, nsamples=301, src_element=1 + 1 * 2000)
    ref = gpu.Sem2dSolver(chunk=2, **kw)
    # the source node is (x, z) = (4, 4): receivers 6 nodes along +x and +z,
    # a corner and two edge nodes
    nodes = [ref.src_node, ref.src_node + 6, ref.src_node + 6 * ref.nx, 0, ref.nx * 20, ref.nx + 20]
    ref.set_receivers(nodes)
    src0, rec0 = ref.run(300)
    cur, prev = ref.current(), ref.previous()
    ref.close()
    f = cur.reshape(ref.nz, ref.nx)
    assert np.isfinite(f).all() and np.abs(f[0, :40]).max() > 0 and np.abs(f[:40, 0]).max() > 0
    assert np.abs(rec0[1:3]).max() > 1e-3 * np.abs(src0).max()
    for chunk, i64 in [(32, False), (1 << 20, False), (32, True)]:
        print("\n[sem2d] 8001^2 chunk %d index64 %d" % (chunk, i64), flush=True)
        G = gpu.Sem2dSolver(chunk=chunk, index64=i64, **kw)
        G.set_receivers(nodes)
        src, rec = G.run(300)
        assert np.array_equal(src, src0) and np.array_equal(rec, rec0)
        assert np.array_equal(G.current(), cur) and np.array_equal(G.previous(), prev)
        G.close()

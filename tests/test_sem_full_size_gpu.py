"""SEM order 4 at the benchmark size (100^3 elements of 20 m, 401^3 nodes),
checked through properties that hold at any size:

* the per-XCD schedule (190 blocks per XCD at this shape, rounds plus pieces)
  and plain plane chunks give the same bits: each node's arithmetic does not
  depend on the tile, block or march that covers it;
* the element kernel with the 8-colour scatter agrees with the separable
  kernel to FP32 summation order (1e-5 relative L2);
* the FP32 loop agrees with the FP64 loop on the same inputs to the 1e-4
  relative L2 of the parity bar (the FP64 loop is pinned to the FP64 oracle
  at small sizes in tests/test_sem64_gpu.py);
* boundary nodes are never written (Dirichlet, SEMsolver.cpp:85-89).

The oracle's element loop is too slow at this size; the small-mesh tests in
tests/test_sem_gpu.py compare the same kernels with it directly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EX, ORDER, H = 100, 4, 20.0
STEPS = 400


def banner(msg):
    print("\n[sem-full] " + msg, flush=True)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def case(gpu):
    lx = H * EX
    model = gpu.sem_model_gaussian(EX, EX, EX, sigma=5.0, vmin=2000.0, vmax=4000.0, seed=20240512)
    c = lx / 2 + 0.5 * H / ORDER
    kw = dict(ex=EX, lx=lx, order=ORDER, dt=2e-4, f0=25.0, model=model, nsamples=STEPS + 1,
              src_point=(c, c, c), weights_ref=False)
    nx = EX * ORDER + 1
    mid = nx // 2
    # receivers on the x line through the centre, 8 to 32 nodes from it (the
    # pulse of a 25 Hz Ricker passes them within the run)
    rec = np.array([ix + mid * nx + mid * nx * nx for ix in (mid - 32, mid - 24, mid - 16, mid - 8,
                                                            mid + 8, mid + 16, mid + 24, mid + 32)], np.int32)
    banner("separable, 16 x 16 tiles, 104-plane chunks, %d steps" % STEPS)
    s = gpu.SemSolver(kernel=gpu.SEM_SEPARABLE, block_waves=16, tile_j=16, chunk_i=104, **kw)
    s.set_receivers(rec)
    src, trc = s.run(STEPS)
    out = dict(kw=kw, rec=rec, nx=nx, src=src, trc=trc, cur=s.current(), prev=s.previous())
    s.close()
    return out


def test_fields_are_live_and_boundary_untouched(case):
    nx = case["nx"]
    f = case["cur"].reshape(nx, nx, nx)
    assert np.isfinite(f).all()
    assert np.abs(case["trc"]).max() > 1e-3 * np.abs(case["src"]).max()
    assert not f[0].any() and not f[-1].any() and not f[:, 0].any() and not f[:, -1].any()
    assert not f[:, :, 0].any() and not f[:, :, -1].any()


def test_per_xcd_schedule_same_bits_as_chunks(gpu, case):
    for waves, tile_j, chunk in [(16, 16, -1), (16, 16, -3)]:
        banner("separable, %d waves x %d rows, schedule %d" % (waves, tile_j, chunk))
        s = gpu.SemSolver(kernel=gpu.SEM_SEPARABLE, block_waves=waves, tile_j=tile_j, chunk_i=chunk, **case["kw"])
        s.set_receivers(case["rec"])
        src, trc = s.run(STEPS)
        assert np.array_equal(src, case["src"])
        assert np.array_equal(trc, case["trc"])
        assert np.array_equal(s.current(), case["cur"])
        assert np.array_equal(s.previous(), case["prev"])
        s.close()


def test_colour_element_kernel_agrees(gpu, case):
    banner("element kernel, 8-colour scatter")
    s = gpu.SemSolver(kernel=gpu.SEM_ELEMENT_COLOR, **case["kw"])
    s.set_receivers(case["rec"])
    src, trc = s.run(STEPS)
    assert rel_l2(s.current(), case["cur"]) < 1e-5
    assert rel_l2(src, case["src"]) < 1e-5
    assert rel_l2(trc, case["trc"]) < 1e-5
    s.close()


def test_fp32_within_parity_bar_of_fp64(gpu, case):
    banner("FP64 loop")
    s = gpu.Sem64Solver(**case["kw"])
    s.set_receivers(case["rec"])
    src, trc = s.run(STEPS)
    f64 = s.field(True)
    s.close()
    assert rel_l2(case["cur"], f64) < 1e-4
    assert rel_l2(case["src"], src) < 1e-4
    assert rel_l2(case["trc"], trc) < 1e-4

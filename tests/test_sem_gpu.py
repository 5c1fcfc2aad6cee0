"""SEM parity: the HIP path (separable and element kernels, through the C ABI)
against the CPU oracle (reference-semantics FP32 element loop and FP64).

The GPU assembles in a different order than the reference's element loop
(FP64 setup, sum-factorised or separable operator), so FP32 results agree to
rounding, not bit for bit.  Tolerances are stated per test; the
floating-point parity bar is 1e-4 relative L2 against the reference."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KERNELS = [0, 1, 2, 3, 4]  # SEM_SEPARABLE, SEM_ELEMENT, SEM_ELEMENT_COLOR, SEM_ELEMENT_MARCH, SEM_ELEMENT_MFMA


def banner(msg):
    print("\n[sem-gpu] " + msg, flush=True)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def pair(gpu, oracle, kernel, ex=6, order=2, lx=240.0, dt=0.002, f0=20.0, nsamples=80,
         model=None, src=None, **kw):
    src = src or (lx / 2 + 1, lx / 2 + 1, lx / 2 + 1)
    s = gpu.SemSolver(ex=ex, lx=lx, order=order, dt=dt, f0=f0, nsamples=nsamples, model=model,
                      src_point=src, kernel=kernel, **kw)
    O = oracle.SemProblem(ex=ex, lx=lx, order=order, dt=dt, f0=f0, nsamples=nsamples, model=model,
                          src_point=src, **{k: v for k, v in kw.items() if k in ("ey", "ez", "ly", "lz")})
    assert np.array_equal(s.rhs, O.rhs)
    assert s.src_element == O.src_element
    return s, O


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_run_matches_reference_semantics(gpu, oracle, kernel, order):
    banner("run kernel=%d order=%d" % (kernel, order))
    dt = {1: 0.003, 2: 0.002, 3: 0.0012, 4: 0.0008, 5: 0.0005}[order]
    s, O = pair(gpu, oracle, kernel, ex=5, order=order, dt=dt, nsamples=70)
    src, _ = s.run(60)
    tr, _ = O.run(60)
    e_field = rel_l2(s.current(), O.current())
    e_prev = rel_l2(s.previous(), O.previous())
    e_tr = rel_l2(src, tr)
    print("rel L2 field %.2e prev %.2e trace %.2e" % (e_field, e_prev, e_tr))
    assert e_field < 1e-5 and e_prev < 1e-5 and e_tr < 1e-5


@pytest.mark.parametrize("kernel", KERNELS)
def test_heterogeneous_model_and_receivers(gpu, oracle, kernel):
    banner("heterogeneous kernel=%d" % kernel)
    rng = np.random.default_rng(11)
    model = (1500 + 2000 * rng.random(7 * 6 * 5)).astype(np.float32)
    s, O = pair(gpu, oracle, kernel, ex=7, ey=6, ez=5, ly=200.0, lz=160.0, order=3, dt=0.0006,
                nsamples=90, model=model, src=(121.0, 81.0, 61.0))
    nodes = np.array([O.src_node, 5, 100, 777, s.nn - 1 - 3 * s.nx, 1234], np.int32)
    s.set_receivers(nodes)
    src, rec = s.run(80)
    tr, orec = O.run(80, receivers=nodes)
    assert rel_l2(s.current(), O.current()) < 1e-5
    assert rel_l2(src, tr) < 1e-5
    for r in range(len(nodes)):
        if np.abs(orec[r]).max() > 0:
            assert rel_l2(rec[r], orec[r]) < 1e-4, r
        else:
            assert np.abs(rec[r]).max() == 0
    assert np.array_equal(rec[0], src)


# Tiles of 64 lanes along x cover k = 0..nx-2 and rows 1..nz-2 (the
# boundary nodes are never updated); meshes whose last tile is nearly empty,
# half full or full along x, and rows that fill the last row tile in part:
#   order 4, ex 17: nx 69, last x tile holds 4 nodes
#   order 4, ex 24: nx 97, 32 nodes
#   order 4, ex 25: nx 101, 36 nodes
#   order 2, ex 40: nx 81, 16 nodes
TILINGS = [(17, 4, 9, 4), (24, 3, 5, 4), (25, 3, 4, 4), (40, 3, 11, 2)]


def test_padded_row_pitch_matches_oracle(gpu, oracle):
    """Order 4, 112 elements along x: 449 nodes per row, whose row pitch
    would be a multiple of 2 KiB and is padded.  After 120 steps this long
    mesh differs from the FP32 oracle by 5.5e-5 (field) and 1.1e-4 (source
    trace) with or without the padding (measured both ways); the check runs
    60 steps at the 1e-4 bar."""
    banner("padded row pitch ex=112")
    ex, ey, ez, h = 112, 3, 4, 20.0
    rng = np.random.default_rng(ex)
    model = (1800 + 1500 * rng.random(ex * ey * ez)).astype(np.float32)
    s, O = pair(gpu, oracle, 0, ex=ex, ey=ey, ez=ez, lx=h * ex, ly=h * ey, lz=h * ez, order=4, dt=0.0004,
                f0=25.0, nsamples=70, model=model, src=(h * (ex // 2) + 1, h * (ey // 2) + 1, h * (ez // 2) + 1))
    src, _ = s.run(60)
    tr, _ = O.run(60)
    assert rel_l2(s.current(), O.current()) < 1e-4
    assert rel_l2(s.previous(), O.previous()) < 1e-4
    assert rel_l2(src, tr) < 1e-4


@pytest.mark.parametrize("ex,ey,ez,order", TILINGS)
def test_partial_tiles_match_oracle(gpu, oracle, ex, ey, ez, order):
    banner("tiling ex=%d ey=%d ez=%d order=%d" % (ex, ey, ez, order))
    h = 20.0
    rng = np.random.default_rng(ex)
    model = (1800 + 1500 * rng.random(ex * ey * ez)).astype(np.float32)
    s, O = pair(gpu, oracle, 0, ex=ex, ey=ey, ez=ez, lx=h * ex, ly=h * ey, lz=h * ez, order=order,
                dt=0.0004, f0=25.0, nsamples=130, model=model,
                src=(h * (ex - 2) + 1, h * (ey // 2) + 1, h * (ez // 2) + 1))
    nx, nz = s.nx, s.nz
    # receivers on the source's row near the source (the source sits in the
    # second-last element, so they lie in the last x tile and the one before)
    sx = O.src_node % nx
    row = O.src_node - sx
    nodes = (row + np.clip(sx + np.array([-8, -4, -2, 2, 4, 6]), 1, nx - 2)).astype(np.int32)
    s.set_receivers(nodes)
    src, rec = s.run(120)
    tr, orec = O.run(120, receivers=nodes)
    assert rel_l2(s.current(), O.current()) < 1e-5
    assert rel_l2(s.previous(), O.previous()) < 1e-5
    assert rel_l2(src, tr) < 1e-5
    for r in range(len(nodes)):
        assert np.abs(orec[r]).max() > 1e-3 * np.abs(tr).max(), r
        assert rel_l2(rec[r], orec[r]) < 1e-5, r
    # boundary nodes are never written
    f = s.current().reshape(s.ny, nz, nx)
    assert not f[:, :, 0].any() and not f[:, :, -1].any() and not f[:, 0, :].any() and not f[:, -1, :].any()


@pytest.mark.parametrize("ex,ey,ez,order", TILINGS[:2])
def test_every_block_shape_gives_the_same_bits(gpu, tmp_path, ex, ey, ez, order):
    # each node's arithmetic is the same whatever tile, block shape or march
    # chunk covers it
    kw = dict(ex=ex, ey=ey, ez=ez, lx=20.0 * ex, ly=20.0 * ey, lz=20.0 * ez, order=order, dt=0.0004,
              f0=25.0, nsamples=40, src_point=(20.0 * (ex - 1) + 1, 21.0, 21.0))
    ref = gpu.SemSolver(**kw)
    ref.run(30)
    # chunk -m: the per-XCD schedules (8 >> (m-1) plane ranges x 2^(m-1) column bands)
    for waves, tile_j, chunk in [(16, 16, 8), (16, 32, 12), (8, 16, 1 << 20), (4, 8, 20), (4, 4, 4),
                                 (8, 8, 16), (16, 16, -1), (4, 8, -1), (16, 16, -2), (4, 8, -3), (16, 16, -4),
                                 (8, 8, -1), (8, 16, -2)]:
        banner("shape waves=%d tile_j=%d chunk=%d" % (waves, tile_j, chunk))
        cache = tmp_path / ("s%d_%d_%d.txt" % (waves, tile_j, chunk))
        cache.write_text("sem %d %d %d %d 0 %d %d %d 1.0\n" % (ex, ey, ez, order, waves, tile_j, chunk))
        s = gpu.SemSolver(**kw)
        assert s.autotune(steps_per_trial=1, cache_path=str(cache)) == (tile_j, waves, chunk)
        s.run(30)
        assert np.array_equal(s.current(), ref.current())
        assert np.array_equal(s.previous(), ref.previous())


@pytest.mark.parametrize("ex,ey,ez,order", TILINGS)
def test_x_remainder_slab_gives_the_same_bits(gpu, ex, ey, ez, order):
    # the last partial x tile marched as a tile (-1) or updated node by node
    # (1; 0 when it holds at most 32 nodes), chunked and per-XCD
    kw = dict(ex=ex, ey=ey, ez=ez, lx=20.0 * ex, ly=20.0 * ey, lz=20.0 * ez, order=order, dt=0.0004,
              f0=25.0, nsamples=40, src_point=(20.0 * (ex - 1) + 1, 21.0, 21.0))
    ref = gpu.SemSolver(k_remainder=-1, **kw)
    src_ref, _ = ref.run(30)
    for mode, chunk in [(1, 0), (0, 0), (1, -1), (0, -3), (1, 12)]:
        banner("x remainder mode %d chunk %d" % (mode, chunk))
        s = gpu.SemSolver(k_remainder=mode, chunk_i=chunk, **kw)
        src, _ = s.run(30)
        assert np.array_equal(src, src_ref)
        assert np.array_equal(s.current(), ref.current())
        assert np.array_equal(s.previous(), ref.previous())
        s.compute_one_step(30, s.i1, s.i2)
        s.close()


def test_xcd_schedule_rounds_and_pieces_bit_exact(gpu):
    # 81 x 61 x 161 nodes, order 2.  With 4 x 4 tiles there are 2 x 40 tile
    # columns: on 32 CUs per XCD, 64 whole-range blocks per XCD and 16 columns
    # cut into 2 pieces; with 16 x 16 tiles 20 columns, each cut into pieces
    kw = dict(ex=40, ey=30, ez=80, lx=800.0, ly=600.0, lz=1600.0, order=2, dt=0.0008, f0=20.0, nsamples=30,
              src_point=(401.0, 301.0, 801.0))
    ref = gpu.SemSolver(block_waves=16, tile_j=16, chunk_i=1 << 20, **kw)
    ref.run(25)
    assert np.abs(ref.current()).max() > 0
    for waves, tile_j, m in [(4, 4, 1), (16, 16, 1), (8, 16, 1), (4, 4, 2), (4, 4, 3), (4, 4, 4), (16, 16, 3)]:
        banner("per-XCD schedule waves=%d tile_j=%d split %d" % (waves, tile_j, m))
        s = gpu.SemSolver(block_waves=waves, tile_j=tile_j, chunk_i=-m, **kw)
        s.run(25)
        assert np.array_equal(s.current(), ref.current())
        assert np.array_equal(s.previous(), ref.previous())
        s.compute_one_step(25, s.i1, s.i2)  # the per-step surface uses the same table
        s.close()
    with pytest.raises(RuntimeError, match="chunk_i"):
        gpu.SemSolver(chunk_i=-5, **kw)


@pytest.mark.parametrize("order", [2, 4, 5])
def test_64bit_offsets_give_the_same_bits(gpu, order):
    # the kernels built for arrays of 4 GiB and more (64-bit offsets), forced
    # on a small mesh: same values as the 32-bit offsets, chunked and per-XCD
    kw = dict(ex=7, ey=5, ez=6, lx=140.0, ly=100.0, lz=120.0, order=order, dt=0.0003, f0=25.0, nsamples=40,
              src_point=(71.0, 51.0, 61.0))
    ref = gpu.SemSolver(**kw)
    src_ref, _ = ref.run(30)
    for chunk in (0, 12, -1):
        s = gpu.SemSolver(index64=True, chunk_i=chunk, **kw)
        src, _ = s.run(30)
        assert np.array_equal(src, src_ref)
        assert np.array_equal(s.current(), ref.current())
        assert np.array_equal(s.previous(), ref.previous())
        s.close()


@pytest.mark.parametrize("order", [2, 4])
def test_mass_and_operator(gpu, oracle, order):
    banner("mass/operator order=%d" % order)
    rng = np.random.default_rng(order)
    model = (1800 + 1500 * rng.random(64)).astype(np.float32)
    O = oracle.SemProblem(ex=4, lx=160.0, order=order, model=model, nsamples=4,
                          src_point=(81.0, 81.0, 81.0))
    u = rng.standard_normal(O.nN).astype(np.float32)
    yo, Mo = O.apply_k(u)
    lst = np.zeros(O.nN, np.int32)
    nI = oracle.lib().orc_sem_interior_nodes(ctypes.byref(O.mesh), lst)
    I = lst[:nI]
    for kernel in KERNELS:
        s = gpu.SemSolver(ex=4, lx=160.0, order=order, dt=1.0, model=model, nsamples=4,
                          src_point=(81.0, 81.0, 81.0), kernel=kernel)
        M = s.mass()
        assert np.abs(M - Mo).max() <= 2e-6 * np.abs(Mo).max()
        y = s.apply_stiffness(u)
        assert rel_l2(y[I], yo[I]) < 2e-5, kernel
        s.close()


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("order", [2, 3, 4])
def test_quadratic_exactness_on_gpu(gpu, oracle, kernel, order):
    s = gpu.SemSolver(ex=4, lx=160.0, order=order, dt=1.0, nsamples=4, src_point=(81.0, 81.0, 81.0),
                      kernel=kernel)
    O = oracle.SemProblem(ex=4, lx=160.0, order=order, nsamples=4, src_point=(81.0, 81.0, 81.0))
    x, yy, z = O.node_coords()
    lst = np.zeros(O.nN, np.int32)
    nI = oracle.lib().orc_sem_interior_nodes(ctypes.byref(O.mesh), lst)
    I = lst[:nI]
    u = ((x - 70) ** 2 + (yy - 90) ** 2 + (z - 75) ** 2).astype(np.float32) / 100.0
    y = s.apply_stiffness(u).astype(np.float64)
    M = s.mass().astype(np.float64) * 1500.0 ** 2
    lap = -y[I] / M[I]
    assert np.allclose(lap, 0.06, rtol=2e-3, atol=2e-5), (lap.min(), lap.max())
    s.close()


@pytest.mark.parametrize("kernel", KERNELS)
def test_per_step_surface_equals_fused_run(gpu, oracle, kernel):
    banner("per-step vs run kernel=%d" % kernel)
    kw = dict(ex=5, lx=200.0, order=2, dt=0.002, f0=20.0, nsamples=40, src_point=(101.0, 101.0, 101.0),
              kernel=kernel)
    a = gpu.SemSolver(**kw)
    b = gpu.SemSolver(**kw)
    for _ in range(30):
        a.step()
    b.run(30)
    assert a.i2 == b.i2
    if kernel == 0:
        # separable kernel: no atomics, the two surfaces give identical floats
        assert np.array_equal(a.current(), b.current())
        assert np.array_equal(a.previous(), b.previous())
    else:
        # element kernel: FP32 atomic scatter, summation order varies by run
        assert rel_l2(a.current(), b.current()) < 1e-6
        assert rel_l2(a.previous(), b.previous()) < 1e-6


@pytest.mark.parametrize("ex,ey,ez,order", [(9, 7, 6, 2), (13, 3, 5, 4), (6, 9, 10, 3), (5, 5, 5, 5)])
def test_element_march_tiles_and_ranges(gpu, oracle, ex, ey, ez, order):
    """The write-once element kernel on meshes whose x and z extents are not
    multiples of its 4-element tiles and whose y layers are cut into ranges:
    within 1e-5 of the oracle, and two runs give the same bits (no atomics)."""
    banner("element march %d x %d x %d order %d" % (ex, ey, ez, order))
    h = 20.0
    rng = np.random.default_rng(ex * 100 + ez)
    model = (1600 + 1800 * rng.random(ex * ey * ez)).astype(np.float32)
    dt = {2: 0.0006, 3: 0.0004, 4: 0.0003, 5: 0.0002}[order]
    src = (h * (ex - 1) + 1, h * (ey // 2) + 1, h * 1 + 1)
    s, O = pair(gpu, oracle, gpu.SEM_ELEMENT_MARCH, ex=ex, ey=ey, ez=ez, lx=h * ex, ly=h * ey, lz=h * ez,
                order=order, dt=dt, f0=25.0, nsamples=80, model=model, src=src)
    src_t, _ = s.run(70)
    tr, _ = O.run(70)
    assert rel_l2(s.current(), O.current()) < 1e-5
    assert rel_l2(s.previous(), O.previous()) < 1e-5
    assert rel_l2(src_t, tr) < 1e-5
    again = gpu.SemSolver(ex=ex, ey=ey, ez=ez, lx=h * ex, ly=h * ey, lz=h * ez, order=order, dt=dt, f0=25.0,
                          nsamples=80, model=model, src_point=src, kernel=gpu.SEM_ELEMENT_MARCH)
    for _ in range(70):
        again.step()
    assert np.array_equal(again.current(), s.current())
    f = s.current().reshape(s.ny, s.nz, s.nx)
    assert not f[:, :, 0].any() and not f[:, :, -1].any() and not f[:, 0, :].any() and not f[0].any()


def test_separable_and_element_kernels_agree(gpu):
    kw = dict(ex=8, lx=320.0, order=4, dt=0.0008, f0=15.0, nsamples=120, src_point=(161.0, 161.0, 161.0))
    a = gpu.SemSolver(kernel=0, **kw)
    b = gpu.SemSolver(kernel=1, **kw)
    ta, _ = a.run(100)
    tb, _ = b.run(100)
    assert rel_l2(a.current(), b.current()) < 1e-5
    assert rel_l2(ta, tb) < 1e-5


@pytest.mark.parametrize("ex,ey,ez", [(7, 5, 6), (4, 4, 4)])
def test_colour_scatter_repeats_bit_for_bit(gpu, ex, ey, ez):
    """8-colour scatter: plain adds in a fixed order, so two runs agree bit
    for bit; the atomic scatter agrees to rounding."""
    banner("colour scatter %dx%dx%d" % (ex, ey, ez))
    kw = dict(ex=ex, ey=ey, ez=ez, lx=40.0 * ex, ly=40.0 * ey, lz=40.0 * ez, order=3, dt=0.0008, f0=15.0,
              nsamples=90, src_point=(20.0 * ex + 1, 20.0 * ey + 1, 20.0 * ez + 1))
    a = gpu.SemSolver(kernel=2, **kw)
    b = gpu.SemSolver(kernel=2, **kw)
    c = gpu.SemSolver(kernel=1, **kw)
    ta, _ = a.run(80)
    for _ in range(80):
        b.step()
    tc, _ = c.run(80)
    assert np.array_equal(a.current(), b.current()) and np.array_equal(a.previous(), b.previous())
    assert rel_l2(a.current(), c.current()) < 2e-6
    assert rel_l2(ta, tc) < 2e-6


def test_fp32_gpu_vs_fp64_oracle(gpu, oracle):
    """Precision of the FP32 time loop against the all-FP64 oracle (exact GLL
    weights, exact geometry)."""
    banner("fp32 gpu vs fp64 oracle")
    kw = dict(ex=8, lx=320.0, order=4, dt=0.0008, f0=15.0, nsamples=160, src_point=(161.0, 161.0, 161.0))
    s = gpu.SemSolver(kernel=0, weights_ref=False, **kw)
    O = oracle.SemProblem(fp64=True, **kw)
    ts, _ = s.run(150)
    to, _ = O.run(150)
    e = rel_l2(s.current(), O.current())
    print("fp32 GPU vs fp64 oracle: field %.2e trace %.2e" % (e, rel_l2(ts, to)))
    assert e < 1e-5 and rel_l2(ts, to) < 1e-5


def test_invalid_inputs(gpu):
    with pytest.raises(gpu.PhcError):
        gpu.SemSolver(ex=3, order=6)
    with pytest.raises(gpu.PhcError):
        gpu.SemSolver(ex=3, model=np.zeros(27, np.float32))
    s = gpu.SemSolver(ex=3, lx=120.0, nsamples=5, src_point=(61.0, 61.0, 61.0))
    with pytest.raises(gpu.PhcError):
        s.compute_one_step(7, 0, 1)
    with pytest.raises(gpu.PhcError):
        s.set_receivers([s.nn])
    s.run(2)
    assert s.check_finite(s.i2) == 0


def test_autotune_keeps_state(gpu, oracle, tmp_path):
    s, O = pair(gpu, oracle, 0, ex=6, order=3, dt=0.001, nsamples=60)
    s.run(20)
    before = s.current().copy()
    cache = str(tmp_path / "tune.txt")
    cfg = s.autotune(steps_per_trial=2, cache_path=cache)
    assert np.array_equal(s.current(), before)
    assert cfg == s.autotune(steps_per_trial=2, cache_path=cache)
    src, _ = s.run(30)
    tr, _ = O.run(50)
    assert rel_l2(s.current(), O.current()) < 1e-5
    assert rel_l2(src, tr[20:]) < 1e-5


@pytest.mark.parametrize("kernel", KERNELS)
def test_source_on_boundary_node(gpu, oracle, kernel):
    """The reference's source point (1001, 1001, 1001) outside a small domain
    selects element 0, whose first node is a Dirichlet boundary node: the
    injection then accumulates in place (SEMsolver.cpp:39-43)."""
    kw = dict(ex=4, lx=160.0, order=2, dt=0.002, f0=20.0, nsamples=60, src_point=(1001.0, 1001.0, 1001.0))
    s = gpu.SemSolver(kernel=kernel, **kw)
    O = oracle.SemProblem(**kw)
    assert O.src_element == 0 and O.src_node == 0
    a = gpu.SemSolver(kernel=kernel, **kw)
    src, _ = s.run(50)
    for _ in range(50):
        a.step()
    tr, _ = O.run(50)
    assert rel_l2(src, tr) < 1e-5
    assert rel_l2(s.current(), O.current()) < 1e-5
    assert rel_l2(a.current(), O.current()) < 1e-5

"""FD parity: the HIP path (through the C ABI) against the CPU oracle.

Reference-rounding mode (ARITH_REF) must be bit-exact with the oracle, which
itself reproduces the reference's published prints.  FP32 mode is held to a
relative L2 of 1e-5 after the stated number of steps (the floating-point
parity bar is 1e-4)."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def make_pair(phc, oracle, nx, ny, nz, L=4, sponge=20, model="two_layer", nsamples=64,
              xs=None, ys=None, zs=None, f0=15.0, dx=10.0, **kw):
    s = phc.FdSetup(nx=nx, ny=ny, nz=nz, L=L, dx=dx, xs=xs, ys=ys, zs=zs, f0=f0,
                    model=model, nsamples=nsamples, sponge_width=sponge)
    P = oracle.FdProblem(nx=nx, ny=ny, nz=nz, L=L, dx=dx, xs=s.xs, ys=s.ys, zs=s.zs, f0=f0,
                         model=s.vp.copy(), nsamples=nsamples, sponge_width=sponge)
    assert np.array_equal(P.rhs, s.rhs) and P.dt == s.dt
    return s, P


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def banner(msg):
    print("\n[fd-gpu] " + msg, flush=True)


@pytest.mark.parametrize("shape,L,sponge", [
    ((40, 40, 40), 4, 6),
    ((37, 29, 70), 4, 8),       # ragged: partial j and k tiles
    ((24, 30, 131), 4, 20),     # grid narrower than twice the sponge on two axes
    ((33, 65, 64), 4, 0),       # no sponge
    ((12, 14, 450), 4, 6),      # row pitch padded (a multiple of 2 KiB otherwise)
])
def test_per_step_surface_bit_exact(gpu, oracle, shape, L, sponge):
    nx, ny, nz = shape
    banner("per-step %s L=%d sponge=%d" % (shape, L, sponge))
    s, P = make_pair(gpu, oracle, nx, ny, nz, L=L, sponge=sponge, nsamples=40)
    sol = gpu.FdSolver(s, arith=gpu.ARITH_REF)
    n = 30
    for it in range(n):
        sol.step(it)
    tr, _ = P.run(n)
    assert np.array_equal(sol.current(), P.current())
    assert np.array_equal(sol.previous(), P.previous())
    assert sol.read_point(sol.i2, s.xs, s.ys, s.zs) == tr[-1]


@pytest.mark.parametrize("L", [1, 2, 3, 4, 5, 6])
def test_radius_bit_exact(gpu, oracle, L):
    banner("radius %d" % L)
    s, P = make_pair(gpu, oracle, 26, 23, 67, L=L, sponge=5, nsamples=30)
    sol = gpu.FdSolver(s, arith=gpu.ARITH_REF)
    src, _ = sol.run(0, 25)
    tr, _ = P.run(25)
    assert np.array_equal(src, tr)
    assert np.array_equal(sol.current(), P.current())
    assert np.array_equal(sol.previous(), P.previous())


@pytest.mark.parametrize("lh", [(4, 3, 2), (2, 4, 1), (6, 1, 3)])
def test_unequal_half_widths_bit_exact(gpu, oracle, lh):
    """-lx/-ly/-lz set per axis (FDTDinit.hpp:96-110; inner3D loops to each
    coef*.size(), FDTDkernels.hpp:16-29): per-step surface and fused run
    bit-exact with the oracle, fields in the reference's per-axis halo
    layout; FP32 mode within 1e-5."""
    lx, ly, lz = lh
    banner("half widths %s" % (lh,))
    s = gpu.FdSetup(nx=30, ny=27, nz=70, L=lx, ly=ly, lz=lz, model="two_layer", nsamples=40, sponge_width=6,
                    f0=20.0)
    P = oracle.FdProblem(nx=30, ny=27, nz=70, L=lx, ly=ly, lz=lz, model=s.vp.copy(), nsamples=40,
                         sponge_width=6, f0=20.0)
    assert P.dt == s.dt and P.coef0 == s.coef0
    sol = gpu.FdSolver(s, arith=gpu.ARITH_REF)
    assert sol.current().shape == P.ext_shape()
    for it in range(10):
        sol.step(it)
    src, _ = sol.run(10, 20)
    tr, _ = P.run(30)
    assert np.array_equal(src, tr[10:])
    assert np.array_equal(sol.current(), P.current())
    assert np.array_equal(sol.previous(), P.previous())
    f32 = gpu.FdSolver(s, arith=gpu.ARITH_FP32)
    f32.run(0, 30)
    assert rel_l2(f32.current(), P.current()) < 1e-5


@pytest.mark.parametrize("nz,mode", [(70, 0), (96, 0), (100, 1), (131, 0), (64, 1), (45, -1)])
def test_k_remainder_slab_bit_exact(gpu, oracle, nz, mode):
    """The last partial 64-point tile along k updated point by point (extra
    blocks of the same launch) gives the bits of the tile march: reference
    rounding = the oracle; FP32 equal between the two paths; with the PML."""
    banner("k remainder nz=%d mode=%d" % (nz, mode))
    s, P = make_pair(gpu, oracle, 33, 29, nz, sponge=7, nsamples=40, xs=16, ys=14, zs=nz - 3)
    sol = gpu.FdSolver(s, arith=gpu.ARITH_REF, k_remainder=mode)
    rec = np.array([[16, 14, nz - 3], [5, 6, nz - 1], [30, 2, 64 if nz > 64 else 1]], np.int32)
    sol.set_receivers(rec)
    src, rv = sol.run(0, 30)
    tr, orv = P.run(30, receivers=rec)
    assert np.array_equal(src, tr) and np.array_equal(rv, orv)
    assert np.array_equal(sol.current(), P.current())
    assert np.array_equal(sol.previous(), P.previous())
    f = [gpu.FdSolver(s, arith=gpu.ARITH_FP32, k_remainder=m) for m in (-1, mode)]
    for x in f:
        x.run(0, 30)
    assert np.array_equal(f[0].current(), f[1].current())
    # PML variant
    s2 = gpu.FdSetup(nx=30, ny=27, nz=nz, L=4, f0=30.0, sponge_width=0, model="two_layer", nsamples=40)
    P2 = oracle.FdProblem(nx=30, ny=27, nz=nz, L=4, f0=30.0, model=s2.vp.copy(), nsamples=40, sponge_width=0,
                          pml=True, ndamp=6)
    p2 = gpu.FdSolver(s2, arith=gpu.ARITH_REF, k_remainder=mode)
    p2.set_pml(ndamp=6, f0=30.0)
    src2, _ = p2.run(0, 30)
    tr2, _ = P2.run(30)
    assert np.array_equal(src2, tr2) and np.array_equal(p2.current(), P2.current())


@pytest.mark.parametrize("ny,nz,mode,tile_j", [(33, 70, 0, 16), (35, 96, 0, 16), (50, 64, 0, 16), (49, 131, 0, 8),
                                              (41, 45, 1, 16)])
def test_j_remainder_slab_bit_exact(gpu, oracle, ny, nz, mode, tile_j):
    """A last j tile holding a quarter of its rows or fewer is updated point
    by point by the same extra blocks as the k slab: reference rounding =
    the oracle, with and without a k slab, in fused runs, in the per-step
    surface and with the PML; FP32 equal to the plain tiles bit for bit."""
    banner("j remainder ny=%d nz=%d mode=%d tile_j=%d" % (ny, nz, mode, tile_j))
    s, P = make_pair(gpu, oracle, 30, ny, nz, sponge=7, nsamples=40, xs=15, ys=ny - 2, zs=nz // 2)
    sol = gpu.FdSolver(s, arith=gpu.ARITH_REF, k_remainder=mode, tile_j=tile_j, block_waves=tile_j // 2)
    rec = np.array([[15, ny - 2, nz // 2], [5, ny - 1, nz - 1], [29, ny - 3, 3]], np.int32)
    sol.set_receivers(rec)
    src, rv = sol.run(0, 25)
    tr, orv = P.run(25, receivers=rec)
    assert np.array_equal(src, tr) and np.array_equal(rv, orv)
    assert np.array_equal(sol.current(), P.current())
    for it in range(25, 30):
        sol.step(it)
    P.run(5)
    assert np.array_equal(sol.current(), P.current())
    assert np.array_equal(sol.previous(), P.previous())
    f = [gpu.FdSolver(s, arith=gpu.ARITH_FP32, k_remainder=m, tile_j=tile_j, block_waves=tile_j // 2)
         for m in (-1, mode)]
    for x in f:
        x.run(0, 30)
    assert np.array_equal(f[0].current(), f[1].current())
    s2 = gpu.FdSetup(nx=28, ny=ny, nz=nz, L=4, f0=30.0, sponge_width=0, model="two_layer", nsamples=30)
    P2 = oracle.FdProblem(nx=28, ny=ny, nz=nz, L=4, f0=30.0, model=s2.vp.copy(), nsamples=30, sponge_width=0,
                          pml=True, ndamp=6)
    p2 = gpu.FdSolver(s2, arith=gpu.ARITH_REF, k_remainder=mode, tile_j=tile_j, block_waves=tile_j // 2)
    p2.set_pml(ndamp=6, f0=30.0)
    src2, _ = p2.run(0, 30)
    tr2, _ = P2.run(30)
    assert np.array_equal(src2, tr2) and np.array_equal(p2.current(), P2.current())


@pytest.mark.parametrize("shape", [(1, 9, 70), (5, 1, 130), (3, 4, 1), (2, 2, 2), (1, 1, 1), (9, 1, 1)])
def test_degenerate_grids_bit_exact(gpu, oracle, shape):
    """Grids one or two points thick along some axes (the stencil then reads
    mostly halo): one partial tile, a march of one plane, a single row or a
    single lane; the fused run and the per-step surface stay bit-exact with
    the oracle."""
    nx, ny, nz = shape
    banner("degenerate grid %s" % (shape,))
    s, P = make_pair(gpu, oracle, nx, ny, nz, sponge=0, nsamples=40)
    sol = gpu.FdSolver(s, arith=gpu.ARITH_REF)
    rec = np.array([[s.xs, s.ys, s.zs], [0, 0, 0], [nx - 1, ny - 1, nz - 1]], np.int32)
    sol.set_receivers(rec)
    src, rv = sol.run(0, 20)
    tr, orv = P.run(20, receivers=rec)
    assert np.abs(tr).max() > 0
    assert np.array_equal(src, tr) and np.array_equal(rv, orv)
    assert np.array_equal(sol.current(), P.current())
    for it in range(20, 25):
        sol.step(it)
    P.run(5)
    assert np.array_equal(sol.current(), P.current())
    assert np.array_equal(sol.previous(), P.previous())


def test_add_rhs_value_equals_add_rhs(gpu):
    """addRHS with the caller's sample (phc_fd_add_rhs_value) gives the bits of
    addRHS from the handle's samples."""
    s = gpu.FdSetup(nx=24, model="two_layer", nsamples=20, sponge_width=4, f0=25.0)
    a = gpu.FdSolver(s, arith=gpu.ARITH_REF)
    b = gpu.FdSolver(s, arith=gpu.ARITH_REF)
    for it in range(12):
        a.step(it)
        gpu.check(gpu.lib().phc_fd_add_rhs_value(b.h, b.i2, s.xs, s.ys, s.zs, float(s.rhs[it])))
        b.compute_one_step_sb(b.i1, b.i2)
        b.i1, b.i2 = b.i2, b.i1
    assert np.array_equal(a.current(), b.current())
    assert np.array_equal(a.previous(), b.previous())
    with pytest.raises(gpu.PhcError):
        gpu.check(gpu.lib().phc_fd_add_rhs_value(b.h, 0, 24, 0, 0, 1.0))
    with pytest.raises(gpu.PhcError):
        gpu.check(gpu.lib().phc_fd_add_rhs_value(b.h, 0, 1, 1, 1, float("nan")))


@pytest.mark.parametrize("tile_j,waves,chunk_i,index64", [
    (8, 4, 7, False), (16, 4, 128, False), (16, 4, 5, True), (8, 4, 64, True),
    (8, 8, 13, False), (16, 8, 256, False), (16, 16, 9, False), (32, 16, 33, True)])
def test_fused_run_bit_exact_with_receivers(gpu, oracle, tile_j, waves, chunk_i, index64):
    banner("fused run tile_j=%d waves=%d chunk_i=%d index64=%d" % (tile_j, waves, chunk_i, index64))
    s, P = make_pair(gpu, oracle, 45, 50, 90, sponge=10, nsamples=80,
                     xs=20, ys=22, zs=41)
    rec = np.array([[20, 22, 41], [0, 0, 0], [44, 49, 89], [10, 30, 64], [20, 22, 40]], np.int32)
    sol = gpu.FdSolver(s, arith=gpu.ARITH_REF, tile_j=tile_j, block_waves=waves, chunk_i=chunk_i,
                       index64=index64)
    sol.set_receivers(rec)
    src, rv = sol.run(0, 60)
    tr, orv = P.run(60, receivers=rec)
    assert np.array_equal(src, tr)
    assert np.array_equal(rv, orv)
    assert np.array_equal(sol.current(), P.current())
    assert np.array_equal(sol.previous(), P.previous())
    # continuing with a second run and then the per-step surface stays exact
    src2, _ = sol.run(60, 10)
    tr2, _ = P.run(10)
    assert np.array_equal(src2, tr2)
    for it in range(70, 75):
        sol.step(it)
    P.run(5)
    assert np.array_equal(sol.current(), P.current())


@pytest.mark.parametrize("split", [1, 2, 3, 4])
@pytest.mark.parametrize("shape,waves,tile_j,krem", [
    ((45, 50, 90), 16, 16, 0),     # few columns: ranges cut into pieces of 1-2 planes
    ((70, 66, 131), 4, 8, 1),      # partial k tile taken point by point by extra blocks
    ((37, 29, 70), 8, 16, -1),
])
def test_xcd_schedule_bit_exact(gpu, oracle, split, shape, waves, tile_j, krem):
    """chunk_i = -m: blocks take (tile column, plane range) from the per-XCD
    table; every point is updated once, with the bits of the oracle, in fused
    runs, in the per-step surface and in the PML step."""
    nx, ny, nz = shape
    banner("xcd schedule split=%d %s waves=%d tile_j=%d krem=%d" % (split, shape, waves, tile_j, krem))
    s, P = make_pair(gpu, oracle, nx, ny, nz, sponge=9, nsamples=50, xs=nx // 2, ys=ny // 3, zs=nz - 5)
    rec = np.array([[nx // 2, ny // 3, nz - 5], [0, 0, 0], [nx - 1, ny - 1, nz - 1], [3, ny - 2, 64]], np.int32)
    sol = gpu.FdSolver(s, arith=gpu.ARITH_REF, block_waves=waves, tile_j=tile_j, chunk_i=-split, k_remainder=krem)
    sol.set_receivers(rec)
    src, rv = sol.run(0, 40)
    tr, orv = P.run(40, receivers=rec)
    assert np.array_equal(src, tr) and np.array_equal(rv, orv)
    assert np.array_equal(sol.current(), P.current())
    assert np.array_equal(sol.previous(), P.previous())
    for it in range(40, 45):
        sol.step(it)
    P.run(5)
    assert np.array_equal(sol.current(), P.current())
    s2 = gpu.FdSetup(nx=nx, ny=ny, nz=nz, L=4, f0=30.0, sponge_width=0, model="two_layer", nsamples=30)
    P2 = oracle.FdProblem(nx=nx, ny=ny, nz=nz, L=4, f0=30.0, model=s2.vp.copy(), nsamples=30, sponge_width=0,
                          pml=True, ndamp=6)
    p2 = gpu.FdSolver(s2, arith=gpu.ARITH_REF, block_waves=waves, tile_j=tile_j, chunk_i=-split,
                      k_remainder=krem)
    p2.set_pml(ndamp=6, f0=30.0)
    src2, _ = p2.run(0, 30)
    tr2, _ = P2.run(30)
    assert np.array_equal(src2, tr2) and np.array_equal(p2.current(), P2.current())


def test_autotune_keeps_state_and_parity(gpu, oracle, tmp_path):
    banner("autotune")
    s, P = make_pair(gpu, oracle, 48, 40, 100, sponge=10, nsamples=40)
    sol = gpu.FdSolver(s, arith=gpu.ARITH_REF)
    sol.run(0, 10)
    before = sol.current().copy()
    cache = str(tmp_path / "tune.txt")
    # a SEM entry (another tag, another field count) ahead of the FD ones
    open(cache, "w").write("sem 10 10 10 4 0 16 16 -1 0.20000\n")
    cfg = sol.autotune(steps_per_trial=3, cache_path=cache)
    assert np.array_equal(sol.current(), before)
    assert cfg[0] % cfg[1] == 0 and (cfg[2] >= 1 or -4 <= cfg[2] <= -1)
    assert cfg == sol.autotune(steps_per_trial=3, cache_path=cache)  # cache hit
    lines = open(cache).read().splitlines()
    assert len(lines) == 2 and lines[1].startswith("fd 48 40 100 4 0 0 0 ")  # requested k mode 0
    src, _ = sol.run(10, 20)
    tr, _ = P.run(30)
    assert np.array_equal(src, tr[10:])
    assert np.array_equal(sol.current(), P.current())


def test_fp32_mode_within_tolerance(gpu, oracle):
    banner("fp32 mode 64^3 x 200")
    s, P = make_pair(gpu, oracle, 64, 64, 64, sponge=12, nsamples=220, f0=25.0)
    sol = gpu.FdSolver(s, arith=gpu.ARITH_FP32)
    src, _ = sol.run(0, 200)
    tr, _ = P.run(200)
    err = rel_l2(sol.current(), P.current())
    print("fp32 rel L2 field %.3e trace %.3e" % (err, rel_l2(src, tr)))
    assert err < 1e-5
    assert rel_l2(src, tr) < 1e-5


def test_published_prints_through_gpu(gpu, oracle):
    """150^3, homogeneous 1500 m/s, f0 = 15 Hz: the GPU run prints what the
    reference's CPU builds printed (test_20240512), and is bit-exact with the
    oracle over the same 401 steps."""
    banner("published 150^3 x 401")
    d = json.load(open(os.path.join(GOLDEN, "published_traces.json")))
    rec = d["result_Cypress_fd_SEQUENTIAL_EPYCROME_01.txt"]
    s = gpu.FdSetup(nx=150, model="homogeneous")
    assert s.nsamples == 1242
    sol = gpu.FdSolver(s, arith=gpu.ARITH_REF)
    src, _ = sol.run(0, 401)
    for step, txt in zip(rec["steps"], rec["values"]):
        if step > 400:
            break
        assert "%f" % src[step] == txt, (step, src[step], txt)
    P = oracle.FdProblem(nx=150, model="homogeneous")
    tr, _ = P.run(401)
    assert np.array_equal(src, tr)
    assert np.array_equal(sol.current(), P.current())


def test_full_size_512_layered_steps(gpu, oracle):
    """The benchmark configuration (512^3, 6 seeded layers, dt from CFL 0.4,
    f0 = 10 Hz): a few reference-rounding steps are bit-exact with the oracle,
    and the FP32 mode stays within 1e-6 of them."""
    banner("512^3 layered, 3 steps")
    s = gpu.FdSetup(nx=512, model=("layered", 6, 20240512), f0=10.0, cfl=0.4, nsamples=1000)
    P = oracle.FdProblem(nx=512, model=s.vp, f0=10.0, dt=s.dt, nsamples=1000)
    sol = gpu.FdSolver(s, arith=gpu.ARITH_REF)
    src, _ = sol.run(0, 3)
    tr, _ = P.run(3)
    assert np.array_equal(src, tr)
    assert np.array_equal(sol.current(), P.current())
    f32 = gpu.FdSolver(s, arith=gpu.ARITH_FP32)
    f32.run(0, 3)
    assert rel_l2(f32.current(), P.current()) < 1e-6


def test_invalid_inputs_raise(gpu):
    s = gpu.FdSetup(nx=16, nsamples=4, sponge_width=4)
    sol = gpu.FdSolver(s)
    with pytest.raises(gpu.PhcError):
        sol.run(2, 5)            # past the end of the source samples
    with pytest.raises(gpu.PhcError):
        sol.compute_one_step_sb(0, 0)
    with pytest.raises(gpu.PhcError):
        sol.set_receivers([[16, 0, 0]])
    s.xs = 99
    with pytest.raises(gpu.PhcError):
        gpu.FdSolver(s)


def test_nonfinite_detection(gpu):
    s = gpu.FdSetup(nx=16, nsamples=8, sponge_width=4)
    sol = gpu.FdSolver(s)
    sol.run(0, 2)
    assert sol.check_finite(sol.i2) == 0
    f = sol.current()
    f[5, 6, 7] = np.nan
    sol.set_field(sol.i2, f)
    assert sol.check_finite(sol.i2) == -4

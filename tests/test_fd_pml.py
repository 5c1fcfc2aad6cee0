"""FD PML boundary (FDTDkernels.hpp:36-104, FDTDutils.hpp:84-171): the CPU
restatement pinned by closed forms and invariants, and the GPU path (through
the C ABI) bit-exact against it in reference rounding."""
import numpy as np
import pytest


def test_profile_closed_form_and_library_helper(phc, oracle):
    n, nd, scale = 40, 7, 0.3
    o = np.zeros(n + 2, np.float32)
    oracle.lib().orc_fd_pml_profile(n, nd, np.float32(scale), o)
    g = np.zeros(n + 2, np.float32)
    phc.check(phc.lib().phc_fd_pml_profile(n, nd, scale, g))
    assert np.array_equal(o, g)
    tmp = np.float32(np.float32(scale) / np.float32(nd * nd))
    for m in range(n + 2):
        if m < nd:
            exp = np.float32((nd - m) ** 2) * tmp
        elif m > n + 1 - nd:
            exp = np.float32((m - (n + 1 - nd)) ** 2) * tmp
        else:
            exp = 0.0
        assert o[m] == exp, m
    assert o[0] == o[n + 1] == np.float32(scale) or abs(o[0] - scale) < 1e-7
    # the reference's defaults: vmin 1500, f0 15, dx 10 -> 16 points
    assert phc.lib().phc_fd_pml_width(1500.0, 15.0, 10.0) == 16
    assert phc.lib().phc_fd_pml_scale(np.float32(8e-4), 4500.0, 16, 10.0) == \
        oracle.lib().orc_fd_pml_scale(np.float32(8e-4), 4500.0, 16, 10.0)


def test_oracle_pml_mirror_symmetry(oracle):
    """Homogeneous model, source at the centre of an odd grid: the field is
    mirror-symmetric bit for bit (a one-off eta index or ghost breaks it)."""
    P = oracle.FdProblem(nx=37, ny=41, nz=39, L=4, f0=30.0, model="homogeneous", nsamples=90, pml=True,
                         sponge_width=0)
    assert list(P.ndamp) == [8, 8, 8]
    P.run(80)
    for c in (P.current(), P.previous()):
        assert np.array_equal(c, c[::-1]) and np.array_equal(c, c[:, ::-1]) and np.array_equal(c, c[:, :, ::-1])
    assert np.max(np.abs(P.phi)) > 0


def test_oracle_pml_width_zero_is_the_plain_loop(oracle):
    kw = dict(nx=24, ny=21, nz=27, L=4, f0=20.0, nsamples=40, sponge_width=0)
    A = oracle.FdProblem(pml=True, ndamp=0, **kw)
    B = oracle.FdProblem(**kw)
    ta, _ = A.run(35)
    tb, _ = B.run(35)
    assert np.array_equal(ta, tb) and np.array_equal(A.current(), B.current())


def test_oracle_pml_absorbs(oracle):
    """A wave leaving through the PML leaves much less energy behind than
    the bare grid (whose halo reflects)."""
    kw = dict(nx=60, L=4, f0=30.0, model="homogeneous", nsamples=700, sponge_width=0)
    A = oracle.FdProblem(pml=True, **kw)
    B = oracle.FdProblem(**kw)
    A.run(650)
    B.run(650)
    nd = int(A.ndamp[0])
    sl = slice(4 + nd, 4 + 60 - nd)
    ea = float(np.sum(A.current()[sl, sl, sl].astype(np.float64) ** 2))
    eb = float(np.sum(B.current()[sl, sl, sl].astype(np.float64) ** 2))
    assert ea < 0.05 * eb, (ea, eb)


def pair(phc, oracle, shape, f0=30.0, xs=None, ys=None, zs=None, nsamples=60, ndamp=None, model="two_layer", L=4,
         ly=None, lz=None):
    nx, ny, nz = shape
    s = phc.FdSetup(nx=nx, ny=ny, nz=nz, L=L, f0=f0, xs=xs, ys=ys, zs=zs, sponge_width=0, model=model,
                    nsamples=nsamples, ly=ly, lz=lz)
    P = oracle.FdProblem(nx=nx, ny=ny, nz=nz, L=L, xs=s.xs, ys=s.ys, zs=s.zs, f0=f0, model=s.vp.copy(),
                         nsamples=nsamples, sponge_width=0, pml=True, ndamp=ndamp, ly=ly, lz=lz)
    assert np.array_equal(P.rhs, s.rhs)
    return s, P


@pytest.mark.gpu
@pytest.mark.parametrize("shape,ndamp,src", [
    ((40, 37, 70), None, None),          # default width (8 at f0 = 30)
    ((33, 29, 45), 5, (2, 14, 22)),      # source inside an i-slab
    ((26, 31, 24), 3, (13, 15, 22)),     # source inside a k-slab
    ((48, 20, 130), 9, None),
])
def test_gpu_pml_bit_exact(gpu, oracle, shape, ndamp, src):
    print("\n[pml] %s ndamp %s src %s" % (shape, ndamp, src), flush=True)
    xs, ys, zs = src or (None, None, None)
    s, P = pair(gpu, oracle, shape, xs=xs, ys=ys, zs=zs, ndamp=ndamp)
    sol = gpu.FdSolver(s, arith=gpu.ARITH_REF)
    nd = sol.set_pml(ndamp=int(P.ndamp[0]), f0=s.f0)
    assert nd == int(P.ndamp[0])
    for it in range(25):
        sol.step(it)
    tr, _ = P.run(25)
    assert np.array_equal(sol.current(), P.current())
    assert np.array_equal(sol.previous(), P.previous())
    assert sol.read_point(sol.i2, s.xs, s.ys, s.zs) == tr[-1]
    # fused, graph-captured run with receivers, continuing
    rec = np.array([[s.xs, s.ys, s.zs], [0, 0, 0], [shape[0] - 1, 3, shape[2] // 2]], np.int32)
    sol.set_receivers(rec)
    src_t, rv = sol.run(25, 20)
    otr, orv = P.run(20, receivers=rec)
    assert np.array_equal(src_t, otr) and np.array_equal(rv, orv)
    assert np.array_equal(sol.current(), P.current())


@pytest.mark.gpu
@pytest.mark.parametrize("chunk,waves", [(-1, 0), (-3, 0), (-4, 0), (8, 0), (-2, 8), (6, 16)])
def test_gpu_pml_bit_exact_schedules(gpu, oracle, chunk, waves):
    # the PML march under the per-XCD schedule (chunk < 0) and under short
    # plane chunks, with block shapes of 8 and 16 waves, on a grid whose tile
    # columns lie inside the inner box, straddle a slab edge, or sit in a
    # slab: bit-identical to the oracle's computeOneStepPML
    print("\n[pml] schedule chunk %d waves %d" % (chunk, waves), flush=True)
    s, P = pair(gpu, oracle, (48, 40, 140), xs=10, ys=20, zs=70, ndamp=6, nsamples=60)
    sol = gpu.FdSolver(s, arith=gpu.ARITH_REF, chunk_i=chunk, block_waves=waves)
    assert sol.set_pml(ndamp=6, f0=s.f0) == 6
    for it in range(10):
        sol.step(it)
    tr, _ = P.run(10)
    assert np.array_equal(sol.current(), P.current())
    rec = np.array([[s.xs, s.ys, s.zs], [1, 1, 1], [47, 39, 139], [24, 20, 135]], np.int32)
    sol.set_receivers(rec)
    src_t, rv = sol.run(10, 45)
    otr, orv = P.run(45, receivers=rec)
    # the pulse has reached the i-slab's points (stored planes 4..9: i = 0..5)
    assert np.abs(otr).max() > 0 and np.abs(P.current()[4:10] - 1e-6).max() > 1.0
    assert np.array_equal(src_t, otr) and np.array_equal(rv, orv)
    assert np.array_equal(sol.current(), P.current())
    assert np.array_equal(sol.previous(), P.previous())


@pytest.mark.gpu
@pytest.mark.parametrize("L,index64,ly,lz", [(1, False, None, None), (2, False, None, None),
                                             (3, False, None, None), (5, False, None, None),
                                             (6, False, None, None), (4, True, None, None),
                                             (4, False, 3, 2), (2, False, 5, 3)])
def test_gpu_pml_bit_exact_other_orders(gpu, oracle, L, index64, ly, lz):
    # -usePML with the stencil of half width L (the march's ring, halo and the
    # pml3D first differences all change with it), also with other widths
    # along y and z (-ly, -lz); and the 64-bit-offset kernel (used for arrays
    # of 4 GiB and more) forced on a small grid
    print("\n[pml] half widths %d %s %s index64 %d" % (L, ly, lz, index64), flush=True)
    s, P = pair(gpu, oracle, (30, 27, 41), nsamples=50, L=L, ly=ly, lz=lz)
    sol = gpu.FdSolver(s, arith=gpu.ARITH_REF, index64=index64)
    nd = sol.set_pml(ndamp=int(P.ndamp[0]), f0=s.f0)
    assert nd == int(P.ndamp[0])
    src_t, _ = sol.run(0, 45)
    tr, _ = P.run(45)
    assert np.abs(tr).max() > 0
    assert np.array_equal(src_t, tr)
    assert np.array_equal(sol.current(), P.current())
    assert np.array_equal(sol.previous(), P.previous())


@pytest.mark.gpu
def test_gpu_pml_fp32_close_and_state_errors(gpu, oracle):
    s, P = pair(gpu, oracle, (44, 41, 66))
    sol = gpu.FdSolver(s, arith=gpu.ARITH_FP32)
    with pytest.raises(gpu.PhcError):
        sol.compute_one_step_pml(0, 1)          # no PML set yet
    sol.set_pml(f0=s.f0)
    with pytest.raises(gpu.PhcError):
        sol.compute_one_step_sb(0, 1)           # a PML is set
    src, _ = sol.run(0, 40)
    tr, _ = P.run(40)
    rel = np.linalg.norm(src.astype(np.float64) - tr) / np.linalg.norm(tr.astype(np.float64))
    assert rel < 1e-5
    sol.set_pml(ndamp=0)                         # removing it restores the plain step
    sol.compute_one_step_sb(0, 1)


@pytest.mark.gpu
def test_gpu_pml_full_size_512(gpu, oracle):
    """The PML benchmark grid (512^3, 6 seeded layers, dt from CFL 0.4, f0 10 Hz,
    the reference's default width: 24 points), with the source moved next to
    the PML corner so that the pulse enters three slabs within the run: 8
    reference-rounding steps bit-exact with the oracle (fields, phi-driven
    slab updates, source trace); the FP32 mode within 1e-5."""
    print("\n[pml] 512^3 layered, source at (30, 30, 30), 8 steps", flush=True)
    s = gpu.FdSetup(nx=512, model=("layered", 6, 20240512), f0=10.0, cfl=0.4, nsamples=1000, sponge_width=0,
                    xs=30, ys=30, zs=30)
    P = oracle.FdProblem(nx=512, xs=30, ys=30, zs=30, f0=10.0, model=s.vp.copy(), dt=s.dt, nsamples=1000,
                         sponge_width=0, pml=True)
    assert int(P.ndamp[0]) == 24
    sol = gpu.FdSolver(s, arith=gpu.ARITH_REF)
    assert sol.set_pml(f0=10.0) == 24
    src_t, _ = sol.run(0, 8)
    tr, _ = P.run(8)
    assert np.abs(tr).max() > 0
    assert np.array_equal(src_t, tr)
    assert np.array_equal(sol.current(), P.current())
    assert np.array_equal(sol.previous(), P.previous())
    sol.close()
    f32 = gpu.FdSolver(s, arith=gpu.ARITH_FP32)
    f32.set_pml(f0=10.0)
    f32.run(0, 8)
    cur = P.current().astype(np.float64)
    assert np.linalg.norm(f32.current() - cur) / np.linalg.norm(cur) < 1e-5

"""Batched multi-shot FD parity: every shot of a batch, advanced by the
persistent batched kernel, against the same shot run alone.

In reference-rounding mode each shot must be bit-exact with the CPU oracle
(which reproduces the reference's published prints); in FP32 mode bit-exact
with the single-grid HIP path, since both run the same tile body on the same
inputs.  Shots are ragged on purpose (partial j and k tiles, grids narrower
than twice the sponge, sources off centre)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = [(40, 33, 70), (29, 47, 64), (52, 20, 35), (17, 18, 19), (23, 66, 130)]


def banner(msg):
    print("\n[batch-gpu] " + msg, flush=True)


def make_shots(phc, dims=DIMS, sponge=8, nsamples=60, seed=11):
    out = []
    for s, (nx, ny, nz) in enumerate(dims):
        out.append(phc.FdSetup(nx, ny, nz, L=4, xs=nx // 3, ys=(2 * ny) // 3, zs=nz // 2 + 1,
                               f0=6.0 + 2.0 * s, cfl=0.4, sponge_width=sponge,
                               model=("layered", 3, seed + s), nsamples=nsamples))
    return out


def receivers(setups):
    rec = []
    for s in setups:
        rec.append([[s.xs, s.ys, s.zs], [0, 0, 0], [s.nx - 1, s.ny - 1, s.nz - 1],
                    [s.nx // 2, s.ny // 2, s.nz - 2]])
    return np.array(rec, np.int32)


def oracle_of(oracle, s):
    P = oracle.FdProblem(nx=s.nx, ny=s.ny, nz=s.nz, L=4, xs=s.xs, ys=s.ys, zs=s.zs, f0=s.f0,
                         cfl=0.4, model=s.vp.copy(), nsamples=s.nsamples, sponge_width=s.sponge_width)
    assert P.dt == s.dt and np.array_equal(P.rhs, s.rhs)
    return P


@pytest.mark.parametrize("krem,chunk", [(0, 16), (-1, 16), (1, 16), (0, -1), (1, -3), (-1, -4)])
def test_ref_mode_bit_exact_with_oracle(gpu, oracle, krem, chunk):
    # krem: shots whose last k tile is partial (nz 70, 130; nz 35 has no full tile)
    # update it point by point (0: when it holds <= 32 points) or march it
    # chunk -m: every shot's per-XCD schedule, XCD x's entries in queue x
    banner("ref mode vs oracle, %d shots, k remainder mode %d, chunk %d" % (len(DIMS), krem, chunk))
    setups = make_shots(gpu)
    rec = receivers(setups)
    b = gpu.FdBatch(setups, arith=gpu.ARITH_REF, chunk_i=chunk, k_remainder=krem)
    b.set_receivers(rec)
    n = 45
    src, rv = b.run(n)
    for k, s in enumerate(setups):
        P = oracle_of(oracle, s)
        tr, orv = P.run(n, receivers=rec[k])
        assert np.array_equal(src[k], tr), "shot %d source trace" % k
        assert np.array_equal(rv[k], orv), "shot %d receiver traces" % k
        assert np.array_equal(b.field(k, current=True), P.current()), "shot %d current" % k
        assert np.array_equal(b.field(k, current=False), P.previous()), "shot %d previous" % k


@pytest.mark.parametrize("tile_j,waves,chunk_i,bpc", [
    (16, 16, 64, 1), (16, 8, 5, 2), (8, 4, 32, 4), (16, 4, 7, 1), (16, 16, 1000, 3),
    (16, 16, -1, 1), (16, 8, -2, 2), (8, 4, -3, 4), (16, 4, -4, 1)])
def test_fp32_bit_exact_with_single_grid(gpu, tile_j, waves, chunk_i, bpc):
    banner("fp32 shape tile_j=%d waves=%d chunk_i=%d blocks/CU=%d" % (tile_j, waves, chunk_i, bpc))
    setups = make_shots(gpu, sponge=6, nsamples=40)
    b = gpu.FdBatch(setups, arith=gpu.ARITH_FP32, tile_j=tile_j, block_waves=waves, chunk_i=chunk_i,
                    blocks_per_cu=bpc)
    n = 33
    src, _ = b.run(n)
    for k, s in enumerate(setups):
        one = gpu.FdSolver(s, arith=gpu.ARITH_FP32)
        tr, _ = one.run(0, n)
        assert np.array_equal(src[k], tr), "shot %d" % k
        assert np.array_equal(b.field(k, True), one.current()), "shot %d current" % k
        assert np.array_equal(b.field(k, False), one.previous()), "shot %d previous" % k
        one.close()


def test_continuation_and_autotune_keep_state(gpu):
    banner("continuation + autotune")
    setups = make_shots(gpu, nsamples=50)
    a = gpu.FdBatch(setups)
    sa, _ = a.run(37)
    b = gpu.FdBatch(setups)
    s1, _ = b.run(12)
    b.autotune(steps_per_trial=1)   # trials run on a copy: the shots must not move
    s2, _ = b.run(25)
    assert np.array_equal(np.concatenate([s1, s2], axis=1), sa)
    for k in range(len(setups)):
        assert np.array_equal(a.field(k, True), b.field(k, True))
        assert np.array_equal(a.field(k, False), b.field(k, False))


def test_graph_replay_matches_eager(gpu):
    """A cached graph (same it0, length, level parity) replayed several times,
    against the timed (eager, event-bracketed) path on a twin batch."""
    banner("graph replay vs eager")
    setups = make_shots(gpu, dims=DIMS[:3], nsamples=20)
    g = gpu.FdBatch(setups)
    e = gpu.FdBatch(setups)
    e.set_timing(True)
    for rep in range(4):
        sg, _ = g.run(10, it0=0)      # even length: same parity, graph reused
        se, _ = e.run(10, it0=0)
        assert np.array_equal(sg, se), "replay %d" % rep
    assert e.last_kernel_ms() > 0.0
    for k in range(len(setups)):
        assert np.array_equal(g.field(k, True), e.field(k, True))


@pytest.mark.parametrize("krem,chunk", [(0, 16), (-1, -1), (1, -4)])
def test_degenerate_shots_bit_exact(gpu, oracle, krem, chunk):
    """Shots one or two points thick along some axes, a single-point shot,
    and a two-point k slab, batched with an ordinary shot: every shot
    bit-exact with the oracle (source trace, receivers, both levels)."""
    dims = [(1, 9, 70), (5, 1, 130), (3, 4, 1), (2, 2, 2), (1, 1, 1), (40, 33, 70)]
    banner("degenerate shots, k remainder mode %d, chunk %d" % (krem, chunk))
    setups = [gpu.FdSetup(nx, ny, nz, L=4, xs=nx // 2, ys=ny // 2, zs=nz // 2, f0=10.0 + s, cfl=0.4,
                          sponge_width=0, model="two_layer", nsamples=50) for s, (nx, ny, nz) in enumerate(dims)]
    rec = np.array([[[s.xs, s.ys, s.zs], [0, 0, 0], [s.nx - 1, s.ny - 1, s.nz - 1]] for s in setups], np.int32)
    b = gpu.FdBatch(setups, arith=gpu.ARITH_REF, chunk_i=chunk, k_remainder=krem)
    b.set_receivers(rec)
    n = 40
    src, rv = b.run(n)
    for k, s in enumerate(setups):
        P = oracle_of(oracle, s)
        tr, orv = P.run(n, receivers=rec[k])
        assert np.abs(tr).max() > 0
        assert np.array_equal(src[k], tr), "shot %d source trace" % k
        assert np.array_equal(rv[k], orv), "shot %d receiver traces" % k
        assert np.array_equal(b.field(k, current=True), P.current()), "shot %d current" % k
        assert np.array_equal(b.field(k, current=False), P.previous()), "shot %d previous" % k


def test_source_on_tile_and_grid_edges(gpu, oracle):
    """Sources on the first/last interior index of each axis and on tile seams."""
    banner("edge sources")
    placements = [(0, 0, 0), (20, 31, 63), (10, 16, 64), (20, 32, 1)]
    setups = []
    for s, (xs, ys, zs) in enumerate(placements):
        setups.append(gpu.FdSetup(21, 33, 70, L=4, xs=xs, ys=ys, zs=zs, f0=10.0, cfl=0.4,
                                  sponge_width=5, model=("layered", 2, 5 + s), nsamples=30))
    b = gpu.FdBatch(setups, arith=gpu.ARITH_REF)
    src, _ = b.run(25)
    for k, s in enumerate(setups):
        P = oracle_of(oracle, s)
        tr, _ = P.run(25)
        assert np.array_equal(src[k], tr), "shot %d" % k
        assert np.array_equal(b.field(k, True), P.current()), "shot %d" % k


def test_full_size_survey_every_shot_exact(gpu, oracle):
    # the benchmark survey (64 shots, axes U[128, 384], 14 GB of fields):
    # every shot of the FP32 batch = the single-grid FP32 path bit for bit,
    # and three of its shots batched in reference rounding = the oracle
    setups = gpu.multi_shot_setups(nshots=64, nsteps=10, seed=2024)
    n = 6
    banner("64-shot survey, FP32, %d steps" % n)
    b = gpu.FdBatch(setups, arith=gpu.ARITH_FP32)
    src, _ = b.run(n)
    for k, s in enumerate(setups):
        one = gpu.FdSolver(s, arith=gpu.ARITH_FP32)
        tr, _ = one.run(0, n)
        assert np.array_equal(src[k], tr), "shot %d" % k
        assert np.array_equal(b.field(k, True), one.current()), "shot %d current" % k
        one.close()
    b.close()
    banner("3 survey shots, reference rounding, vs the oracle")
    sub = setups[:3]
    r = gpu.FdBatch(sub, arith=gpu.ARITH_REF)
    rsrc, _ = r.run(n)
    for k, s in enumerate(sub):
        P = oracle.FdProblem(nx=s.nx, ny=s.ny, nz=s.nz, L=4, xs=s.xs, ys=s.ys, zs=s.zs, f0=s.f0, dt=s.dt,
                             model=s.vp.copy(), nsamples=s.nsamples, sponge_width=s.sponge_width)
        assert np.array_equal(P.rhs, s.rhs)
        tr, _ = P.run(n)
        assert np.abs(tr).max() > 0
        assert np.array_equal(rsrc[k], tr), "shot %d" % k
        assert np.array_equal(r.field(k, True), P.current()), "shot %d current" % k
    r.close()

"""Host-side checks of the batched-shot entry points (argument validation and
the survey generator); no device work, so they run without a GPU."""
import ctypes

import numpy as np


def _params(phc, setups, **kw):
    n = len(setups)
    dims = np.array([[s.nx, s.ny, s.nz] for s in setups], np.int32)
    src = np.array([[s.xs, s.ys, s.zs] for s in setups], np.int32)
    vps = (ctypes.c_void_p * n)(*[s.vp.ctypes.data for s in setups])
    rhs = (ctypes.c_void_p * n)(*[s.rhs.ctypes.data for s in setups])
    p = phc.BatchParams()
    p.nshots = n
    p.dims, p.src = dims.ctypes.data, src.ctypes.data
    p.vp, p.rhs = ctypes.cast(vps, ctypes.c_void_p), ctypes.cast(rhs, ctypes.c_void_p)
    p.nsamples = setups[0].nsamples
    p.half_width = 4
    p.coef = setups[0].cx.ctypes.data
    p.coef0 = setups[0].coef0
    p.sponge_width = 8
    for k, v in kw.items():
        setattr(p, k, v)
    return p, (dims, src, vps, rhs)


def _create(phc, p):
    h = ctypes.c_void_p()
    st = phc.lib().phc_batch_create(ctypes.byref(p), ctypes.byref(h))
    return st, phc.lib().phc_last_error().decode()


def test_invalid_batches_are_rejected(phc):
    setups = [phc.FdSetup(20, 21, 22, nsamples=10), phc.FdSetup(18, 30, 25, nsamples=10)]
    cases = [
        (dict(nshots=0), "no shots"),
        (dict(half_width=3), "half width 4"),
        (dict(sponge_width=65), "sponge width"),
        (dict(arith=7), "arithmetic"),
        (dict(nsamples=0), "nsamples"),
        (dict(block_waves=16, tile_j=24), "no batched kernel"),
        (dict(blocks_per_cu=9), "blocks_per_cu"),
    ]
    for kw, msg in cases:
        p, keep = _params(phc, setups, **kw)
        st, err = _create(phc, p)
        assert st == -1 and msg in err, (kw, st, err)
    # source outside its shot, empty shot
    p, keep = _params(phc, setups)
    bad = keep[1].copy()
    bad[1, 1] = 30
    p.src = bad.ctypes.data
    st, err = _create(phc, p)
    assert st == -1 and "shot 1" in err and "outside" in err
    dims = keep[0].copy()
    dims[0, 2] = 0
    p, keep = _params(phc, setups)
    p.dims = dims.ctypes.data
    st, err = _create(phc, p)
    assert st == -1 and "empty" in err
    assert phc.lib().phc_batch_run(None, 0, 1) == -1
    assert phc.lib().phc_batch_destroy(None) == 0


def test_multi_shot_survey_generator(phc):
    setups = phc.multi_shot_setups(nshots=6, nmin=40, nmax=60, nsteps=12, seed=3)
    assert len(setups) == 6
    for s in setups:
        assert 40 <= min(s.nx, s.ny, s.nz) and max(s.nx, s.ny, s.nz) <= 60
        assert 5.0 <= s.f0 <= 15.0 and s.nsamples == 12 and s.sponge_width == 20
        assert s.dt == setups[0].dt and np.array_equal(s.cx, setups[0].cx)
        assert s.vp.min() > 0
    # same seed, same survey
    again = phc.multi_shot_setups(nshots=6, nmin=40, nmax=60, nsteps=12, seed=3)
    assert all(np.array_equal(a.vp, b.vp) and a.f0 == b.f0 for a, b in zip(setups, again))

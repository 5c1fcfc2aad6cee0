"""phc_lb_train_noalloc: training without device allocation (static scratch)
against phc_lb_train, through the C ABI.  Two calls in a row, the second
with a smaller batch and fewer steps; each must leave the same parameters
and loss history bit for bit (training reduces in a fixed order).  Sizes
beyond the scratch limits are refused."""
import ctypes

import numpy as np
import pytest

import lb_oracle as lo

pytestmark = pytest.mark.gpu


def samples(rng, counts):
    out = []
    for n in counts:
        S = np.zeros((n, lo.NSTRIDE), np.float32)
        S[:, :18] = rng.uniform(-1, 1, (n, 18))
        S[:, 18] = rng.uniform(0.1, 0.5, n)
        S[:, 19] = rng.normal(0, 0.5, n)
        S[:, 20] = rng.integers(0, lo.W, n)
        out.append(S)
    return out


def model(gpu, data):
    lb = gpu.LearnedBoundary(seed=21)
    lb.set_capacity(8192)
    for f, S in enumerate(data):
        if len(S):
            lb.add_samples(f, S)
    lb.set_train_noise(0.02)
    return lb


def train(gpu, lb, fn, steps, batch, seed):
    hist = np.zeros(max(steps, 1), np.float32)
    gpu.check(getattr(gpu.lib(), fn)(lb.h, steps, batch, 1e-3, seed, hist.ctypes.data_as(ctypes.c_void_p)))
    return hist[:steps]


def test_train_noalloc_equals_train(gpu):
    print("\n[no-alloc] phc_lb_train_noalloc vs phc_lb_train", flush=True)
    data = samples(np.random.default_rng(17), [5000, 3000, 0, 2600, 700, 8000])
    a, b = model(gpu, data), model(gpu, data)
    for steps, batch, seed in ((30, 8192, 4), (12, 1000, 5)):
        ha = train(gpu, a, "phc_lb_train", steps, batch, seed)
        hb = train(gpu, b, "phc_lb_train_noalloc", steps, batch, seed)
        assert np.all(np.isfinite(ha))
        assert np.array_equal(ha, hb)
        assert np.array_equal(a.params(), b.params())
    # beyond the static array: refused, state unchanged
    before = b.params().copy()
    for steps, batch in ((5, 8193), (16385, 64)):
        with pytest.raises(gpu.PhcError):
            train(gpu, b, "phc_lb_train_noalloc", steps, batch, 1)
    assert np.array_equal(b.params(), before)
    a.close()
    b.close()

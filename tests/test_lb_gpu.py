"""Learned thin boundary on the GPU against the FP64 oracle (oracle/lb_oracle.py):
the training kernels' loss and gradient, the MFMA inference, the collection of
training samples from a wide-sponge run, the correction pass of a thin run,
and a few learned-boundary steps against a CPU emulation."""
import numpy as np
import pytest

import lb_oracle as lo

pytestmark = pytest.mark.gpu


def banner(msg):
    print("\n[lb-gpu] " + msg, flush=True)


def random_samples(rng, counts):
    out = []
    for n in counts:
        S = np.zeros((n, lo.NSTRIDE), np.float32)
        S[:, :18] = rng.uniform(-1, 1, (n, 18))
        S[:, 18] = rng.uniform(0.1, 0.5, n)
        S[:, 19] = rng.normal(0, 0.5, n)
        S[:, 20] = rng.integers(0, lo.W, n)
        out.append(S)
    return out


def test_loss_and_gradient_match_oracle(gpu):
    banner("eval loss/grad vs oracle")
    rng = np.random.default_rng(11)
    lb = gpu.LearnedBoundary(seed=3)
    samples = random_samples(rng, [300, 17, 0, 129, 64, 1])
    lb.set_capacity(512)
    for f, S in enumerate(samples):
        if len(S):
            lb.add_samples(f, S)
    assert list(lb.counts()) == [len(S) for S in samples]
    theta = lb.params()
    loss, g = lb.eval()
    oloss, og = lo.loss_grad(theta, samples)
    assert loss == pytest.approx(oloss, rel=1e-5)
    assert np.linalg.norm(g - og) <= 1e-4 * np.linalg.norm(og)
    for f in range(lo.FACES):
        if len(samples[f]) > 16:
            assert np.linalg.norm(g[f] - og[f]) <= 1e-4 * np.linalg.norm(og[f]), f
    assert np.all(g[2] == 0)


def test_mfma_inference_matches_oracle(gpu):
    banner("MFMA inference vs oracle")
    rng = np.random.default_rng(12)
    lb = gpu.LearnedBoundary(seed=5)
    theta = rng.normal(0, 0.4, (lo.FACES, lo.NP)).astype(np.float32)
    lb.set_params(theta)
    for f, n in ((0, 1), (3, 31), (4, 33), (5, 1000)):
        S = random_samples(rng, [n])[0]
        y = lb.infer(f, S)
        oy, _ = lo.forward(lo.bf16(theta[f]), S[:, :lo.NIN].astype(np.float64),
                           S[:, lo.NIN + 1].astype(int))
        scale = max(np.max(np.abs(oy)), 1e-30)
        assert np.max(np.abs(y - oy)) <= 2e-4 * scale, (f, n, np.max(np.abs(y - oy)) / scale)


def wide_run(gpu, n_wide, sponge, steps, src, seed=2, nl=3):
    s = gpu.FdSetup(n_wide, n_wide, n_wide, L=4, xs=src[0], ys=src[1], zs=src[2], f0=20.0, cfl=0.4,
                    sponge_width=sponge, model=("layered", nl, seed), nsamples=steps + 1)
    sol = gpu.FdSolver(s, arith=gpu.ARITH_FP32, init_value=0.0)
    sol.run(0, steps)
    return s, sol


def test_collected_samples_match_oracle(gpu):
    banner("collection vs oracle")
    s, sol = wide_run(gpu, 56, 10, 70, (28, 26, 30))
    o, n = (15, 14, 16), (26, 25, 24)
    lb = gpu.LearnedBoundary(seed=1)
    lb.set_capacity(20000)
    lb.collect(sol, sol.i2, o, n, halo=0.0, keep=1.0, s_min=0.0, seed=7)
    pn, pm = sol.current().astype(np.float64), sol.previous().astype(np.float64)
    vp3 = s.vp.reshape(s.nx, s.ny, s.nz)
    c1 = float(s.cx[1])
    chat = s.cx.astype(np.float64) / c1
    total = 0
    for f in range(lo.FACES):
        S = lb.samples(f)
        X, sc, d, v, t = lo.layer_inputs(pn, pm, vp3, o, n, f, 0.0, c1, chat)
        keep = sc > lo.S_TINY
        assert len(S) == int(keep.sum()), f
        idx = S[:, 21].astype(int)
        assert len(np.unique(idx)) == len(idx)
        assert np.all(keep[idx])
        assert np.allclose(S[:, :lo.NIN], X[idx], rtol=1e-5, atol=1e-6)
        assert np.allclose(S[:, lo.NIN], t[idx], rtol=1e-4, atol=1e-5)
        assert np.array_equal(S[:, lo.NIN + 1].astype(int), d[idx])
        total += len(S)
    assert total > 1000
    # the store fills in compact order (the same order on every run)
    for f in range(lo.FACES):
        assert np.all(np.diff(lb.samples(f)[:, 21]) > 0), f
    # subsampling keeps about the requested fraction; a second collection of
    # the same state appends an identical block
    lb.clear_samples()
    lb.collect(sol, sol.i2, o, n, keep=0.25, seed=9)
    assert 0.15 * total < lb.counts().sum() < 0.35 * total
    first = [lb.samples(f).copy() for f in range(lo.FACES)]
    lb.collect(sol, sol.i2, o, n, keep=0.25, seed=9)
    for f in range(lo.FACES):
        S = lb.samples(f)
        k = len(first[f])
        assert len(S) == 2 * k and np.array_equal(S[:k], first[f]) and np.array_equal(S[k:], first[f]), f


def thin_solver(gpu, n, steps, seed=4, src=None):
    src = src or (n // 2, n // 2 - 1, n // 2 + 1)
    s = gpu.FdSetup(n, n, n, L=4, xs=src[0], ys=src[1], zs=src[2], f0=25.0, cfl=0.4, sponge_width=0,
                    model=("layered", 3, seed), nsamples=steps + 1)
    return s


def chat_of(s):
    """c_l / c_1 in float32, as the FD handle forms them for the layer."""
    c = np.asarray(s.cx, np.float32)
    return (c / c[1]).astype(np.float32)


def test_correction_pass_matches_oracle(gpu):
    banner("correction pass vs oracle")
    rng = np.random.default_rng(13)
    s = thin_solver(gpu, 40, 60)
    sol = gpu.FdSolver(s, arith=gpu.ARITH_FP32, init_value=0.0)
    sol.run(0, 45)           # a wavefield that reaches the layer
    lb = gpu.LearnedBoundary(seed=2)
    theta = rng.normal(0, 0.3, (lo.FACES, lo.NP)).astype(np.float32)
    lb.set_params(theta)
    sol.set_learned_boundary(lb, halo=0.0)
    got = sol.lb_corrections()
    pn, pm = sol.current().astype(np.float64), sol.previous().astype(np.float64)
    vp3 = s.vp.reshape(s.nx, s.ny, s.nz)
    ref = lo.corrections(theta, pn, pm, vp3, (s.nx, s.ny, s.nz), 0.0, float(s.cx[1]), chat_of(s))
    for f in range(lo.FACES):
        scale = max(np.max(np.abs(ref[f])), 1e-30)
        assert np.max(np.abs(got[f] - ref[f])) <= 3e-4 * scale, f
        assert np.max(np.abs(ref[f])) > 0


def test_learned_steps_match_cpu_emulation(gpu, oracle):
    """REF arithmetic on both sides: the stencil is bit-exact, the corrections
    agree to the MFMA inference tolerance."""
    banner("learned steps vs CPU emulation")
    rng = np.random.default_rng(14)
    n = 36
    s = thin_solver(gpu, n, 80, src=(18, 17, 19))
    theta = (rng.normal(0, 0.3, (lo.FACES, lo.NP)) * 0.5).astype(np.float32)
    lb = gpu.LearnedBoundary(seed=2)
    lb.set_params(theta)
    P = oracle.FdProblem(nx=n, L=4, xs=s.xs, ys=s.ys, zs=s.zs, f0=s.f0, cfl=0.4, model=s.vp.copy(),
                         nsamples=s.nsamples, sponge_width=0)
    assert np.array_equal(P.rhs, s.rhs)
    sol = gpu.FdSolver(s, arith=gpu.ARITH_REF, init_value=1e-6)
    sol.set_learned_boundary(lb)   # halo = init_value
    vp3 = s.vp.reshape(n, n, n)
    c1 = float(s.cx[1])
    nsteps = 70
    for it in range(nsteps):
        corr = lo.corrections(theta, P.current().astype(np.float64), P.previous().astype(np.float64), vp3,
                              (n, n, n), np.float64(np.float32(1e-6)), c1, chat_of(s))
        P.run(1)
        cur = P.current()
        upd = cur.astype(np.float64)
        lo.apply_corrections(upd, corr, (n, n, n))
        cur[...] = upd.astype(np.float32)
        sol.step(it)
    g, r = sol.current(), P.current()
    rel = np.linalg.norm((g - r).astype(np.float64)) / np.linalg.norm(r.astype(np.float64) - 1e-6)
    assert rel < 1e-4, rel
    # the fused, graph-captured run gives the same as the per-step surface
    sol2 = gpu.FdSolver(s, arith=gpu.ARITH_REF, init_value=1e-6)
    sol2.set_learned_boundary(lb)
    sol2.run(0, nsteps)
    assert np.array_equal(sol2.current(), g)
    assert np.array_equal(sol2.previous(), sol.previous())


def test_learned_steps_full_size_264(gpu, oracle):
    """The learned grid of the bench (256^3 region + 4-point layer = 264^3):
    from a smooth field that fills the layers, 3 learned steps (reference
    rounding) against the CPU emulation -- stencil bit-exact, corrections
    to the MFMA inference tolerance."""
    banner("learned steps at 264^3 vs CPU emulation")
    n = 264
    rng = np.random.default_rng(15)
    s = thin_solver(gpu, n, 8, src=(130, 131, 129))
    theta = (rng.normal(0, 0.3, (lo.FACES, lo.NP)) * 0.5).astype(np.float32)
    lb = gpu.LearnedBoundary(seed=2)
    lb.set_params(theta)
    P = oracle.FdProblem(nx=n, L=4, xs=s.xs, ys=s.ys, zs=s.zs, f0=s.f0, cfl=0.4, model=s.vp.copy(),
                         nsamples=s.nsamples, sponge_width=0)
    x = np.arange(n + 8, dtype=np.float64)
    f1 = (np.sin(0.043 * x)[:, None, None] * np.cos(0.051 * x)[None, :, None] *
          np.sin(0.037 * x + 0.3)[None, None, :]).astype(np.float32)
    f1[:4] = f1[-4:] = 0.0        # the halo holds the constant h = 0
    f1[:, :4] = f1[:, -4:] = 0.0
    f1[:, :, :4] = f1[:, :, -4:] = 0.0
    f0 = (0.98 * f1).astype(np.float32)
    P.p1[:] = f1.reshape(-1)      # current (i2)
    P.p0[:] = f0.reshape(-1)      # previous (i1)
    sol = gpu.FdSolver(s, arith=gpu.ARITH_REF, init_value=0.0)
    sol.set_field(sol.i2, f1)
    sol.set_field(sol.i1, f0)
    sol.set_learned_boundary(lb, halo=0.0)
    vp3 = s.vp.reshape(n, n, n)
    c1 = float(s.cx[1])
    for it in range(3):
        corr = lo.corrections(theta, P.current().astype(np.float64), P.previous().astype(np.float64), vp3,
                              (n, n, n), 0.0, c1, chat_of(s))
        assert max(np.abs(c).max() for c in corr) > 0
        P.run(1)
        cur = P.current()
        upd = cur.astype(np.float64)
        lo.apply_corrections(upd, corr, (n, n, n))
        cur[...] = upd.astype(np.float32)
        sol.step(it)
    g, r = sol.current(), P.current()
    rel = np.linalg.norm((g - r).astype(np.float64)) / np.linalg.norm(r.astype(np.float64))
    print("[lb-gpu] 264^3 learned steps rel L2 %.3e" % rel, flush=True)
    assert rel < 1e-4, rel
    # the fused, graph-captured run gives the per-step surface bit for bit
    sol2 = gpu.FdSolver(s, arith=gpu.ARITH_REF, init_value=0.0)
    sol2.set_field(sol2.i2, f1)
    sol2.set_field(sol2.i1, f0)
    sol2.set_learned_boundary(lb, halo=0.0)
    sol2.run(0, 3)
    assert np.array_equal(sol2.current(), g)
    assert np.array_equal(sol2.previous(), sol.previous())


def test_training_reduces_the_loss(gpu):
    banner("training")
    s, sol = wide_run(gpu, 56, 10, 90, (30, 27, 26), seed=3)
    lb = gpu.LearnedBoundary(seed=4)
    lb.set_capacity(200000)
    lb.collect(sol, sol.i2, (15, 14, 16), (26, 25, 24), s_min=1e-8)
    l0, _ = lb.eval(with_grad=False)
    hist = lb.train(300, batch_per_face=2048, lr=2e-3, seed=1)
    l1, _ = lb.eval(with_grad=False)
    assert np.all(np.isfinite(hist))
    assert l1 < 0.5 * l0, (l0, l1)


def test_onpolicy_samples_match_oracle(gpu):
    banner("on-policy collection vs oracle")
    import learned
    rng = np.random.default_rng(21)
    region, margin = 32, 6
    o = margin + 10
    N = region + 2 * o
    dt = learned.fd_dt()
    inter, vel = learned.random_layers(rng, region)
    src = (o + 15, o + 13, o + 17)
    ws = gpu.FdSetup(N, N, N, L=4, xs=src[0], ys=src[1], zs=src[2], f0=20.0, cfl=0.4, sponge_width=10,
                     model=learned.layered_vp2dt2((N, N, N), o, inter, vel, dt), nsamples=120)
    ts = gpu.FdSetup(region, region, region, L=4, xs=15, ys=13, zs=17, f0=20.0, cfl=0.4, sponge_width=0,
                     model=learned.layered_vp2dt2((region,) * 3, 0, inter, vel, dt), nsamples=120)
    wide = gpu.FdSolver(ws, arith=gpu.ARITH_FP32, init_value=0.0)
    thin = gpu.FdSolver(ts, arith=gpu.ARITH_FP32, init_value=0.0)
    lb = gpu.LearnedBoundary(seed=6)
    thin.set_learned_boundary(lb, halo=0.0)
    for it in range(70):
        wide.step(it)
        thin.step(it)
    lb.set_capacity(100000)
    lb.collect_onpolicy(thin, wide, (o, o, o), halo=0.0)
    tpn, tpm = thin.current().astype(np.float64), thin.previous().astype(np.float64)
    wc = wide.current().astype(np.float64)
    vp3 = ts.vp.reshape(region, region, region)
    c1 = float(ts.cx[1])
    chat = ts.cx.astype(np.float64) / c1
    total = 0
    for f in range(lo.FACES):
        S = lb.samples(f)
        X, sc, d, t, ok = lo.onpolicy_targets(tpn, tpm, vp3, wc, (o, o, o), (region,) * 3, f, 0.0, chat, c1)
        assert len(S) == int(ok.sum()), f
        idx = S[:, 21].astype(int)
        assert np.all(ok[idx])
        assert np.allclose(S[:, :lo.NIN], X[idx], rtol=1e-5, atol=1e-6)
        assert np.allclose(S[:, lo.NIN], t[idx], rtol=1e-4, atol=1e-5)
        assert np.array_equal(S[:, lo.NIN + 1].astype(int), d[idx])
        total += len(S)
    assert total > 1000


@pytest.mark.parametrize("noise", [0.0, 0.02])
def test_training_repeats_bit_for_bit(gpu, noise):
    """Per-block gradient sums added in a fixed order: two trainings from the
    same parameters, samples and seed end with identical parameters and loss
    histories (with and without input noise)."""
    banner("training determinism, noise %g" % noise)
    rng = np.random.default_rng(5)
    samples = random_samples(rng, [3000, 2500, 0, 1800, 900, 4000])
    runs = []
    for _ in range(2):
        lb = gpu.LearnedBoundary(seed=9)
        lb.set_capacity(4096)
        for f, S in enumerate(samples):
            if len(S):
                lb.add_samples(f, S)
        lb.set_train_noise(noise)
        hist = lb.train(40, batch_per_face=1024, lr=1e-3, seed=3)
        runs.append((hist, lb.params()))
        lb.close()
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
    assert np.all(np.isfinite(runs[0][0]))

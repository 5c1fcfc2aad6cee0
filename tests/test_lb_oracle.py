"""CPU checks of the learned-boundary oracle (oracle/lb_oracle.py): the
analytic gradient against central finite differences, the layer geometry, and
the training target against the FD operator it stands for."""
import numpy as np
import pytest

import lb_oracle as lo


def random_samples(rng, counts):
    out = []
    for n in counts:
        S = np.zeros((n, lo.NSTRIDE))
        S[:, :18] = rng.uniform(-1, 1, (n, 18))
        S[:, 18] = rng.uniform(0.1, 0.5, n)
        S[:, 19] = rng.normal(0, 0.5, n)
        S[:, 20] = rng.integers(0, lo.W, n)
        out.append(S)
    return out


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    samples = random_samples(rng, [17, 0, 9, 23, 5, 11])
    theta = rng.normal(0, 0.3, (lo.FACES, lo.NP))
    g = lo.loss_grad64(theta, samples)
    idx = rng.choice(lo.FACES * lo.NP, 300, replace=False)
    eps = 1e-6
    num = np.zeros(len(idx))
    for q, i in enumerate(idx):
        tp = theta.copy().reshape(-1)
        tm = theta.copy().reshape(-1)
        tp[i] += eps
        tm[i] -= eps
        num[q] = (lo.loss_at(tp.reshape(lo.FACES, lo.NP), samples) -
                  lo.loss_at(tm.reshape(lo.FACES, lo.NP), samples)) / (2 * eps)
    ana = g.reshape(-1)[idx]
    assert np.linalg.norm(ana - num) <= 1e-6 * np.linalg.norm(num)
    # every parameter group is exercised
    assert np.all(np.abs(g[0, lo.OFF_W1:lo.OFF_B1]) > 0)
    assert np.all(g[1] == 0)  # face without samples


def test_loss_grad_at_bf16_equals_fp64_formula_at_rounded_weights():
    rng = np.random.default_rng(4)
    samples = random_samples(rng, [8] * 6)
    theta = rng.normal(0, 0.3, (lo.FACES, lo.NP)).astype(np.float32)
    loss, g = lo.loss_grad(theta, samples)
    r = lo.bf16(theta).reshape(lo.FACES, lo.NP)
    assert loss == pytest.approx(lo.loss_at(r, samples), rel=1e-12)
    assert np.allclose(g, lo.loss_grad64(r, samples), rtol=1e-12, atol=1e-15)
    # bf16 rounding: 8 significant bits, nearest even
    assert lo.bf16(np.float32(1.0 + 2 ** -8)) == 1.0
    assert lo.bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6
    assert lo.bf16(np.float32(-3.0)) == -3.0


def test_layer_geometry_covers_the_union_once_per_axis():
    n = (10, 9, 12)
    seen = np.zeros(n, int)
    for f in range(lo.FACES):
        c, d = lo.face_points(n, f)
        a, side, nu, nv = lo.face_geom(n, f)
        assert len(c) == lo.W * nu * nv
        assert np.all(d == np.where(side, n[a] - 1 - c[:, a], c[:, a]))
        np.add.at(seen, (c[:, 0], c[:, 1], c[:, 2]), 1)
    inl = [np.minimum(np.arange(m), m - 1 - np.arange(m)) < lo.W for m in n]
    expect = (inl[0][:, None, None].astype(int) + inl[1][None, :, None] + inl[2][None, None, :])
    assert np.array_equal(seen, expect)


def _laplacian_update(p, c, R, pts):
    """sum over axes and l of c_l (p(P+l e) + p(P-l e)) at extended indices pts."""
    out = np.zeros(len(pts))
    for a in range(3):
        e = np.zeros(3, int)
        e[a] = 1
        for l in range(1, R + 1):
            qp = pts + l * e
            qm = pts - l * e
            out += c[l] * (p[qp[:, 0], qp[:, 1], qp[:, 2]] + p[qm[:, 0], qm[:, 1], qm[:, 2]])
    return out


def test_target_closes_the_truncated_stencil():
    """With y equal to the target, the thin grid's stencil plus the correction
    equals the wide grid's stencil at every layer point (edges and corners
    included, where two or three faces add their terms)."""
    rng = np.random.default_rng(5)
    R = lo.R
    c = np.array([0.0, 1.6, -0.2, 0.0253968, -0.0017857]) / 100.0
    c1 = c[1]
    chat = c / c1
    h = 1e-6
    n = (11, 10, 13)
    o = (6, 7, 5)
    N = tuple(o[a] + n[a] + 6 for a in range(3))
    wide = h + rng.normal(0, 1, tuple(m + 2 * R for m in N))
    prev = h + rng.normal(0, 1, wide.shape)
    vp = rng.uniform(0.01, 0.05, N)
    # the thin grid: the region of the wide grid, halo = h
    thin = np.full(tuple(m + 2 * R for m in n), h)
    thin[R:R + n[0], R:R + n[1], R:R + n[2]] = wide[o[0] + R:o[0] + R + n[0], o[1] + R:o[1] + R + n[1],
                                                  o[2] + R:o[2] + R + n[2]]
    corr = np.zeros(n)
    for f in range(lo.FACES):
        X, s, d, v, t = lo.layer_inputs(wide, prev, vp, o, n, f, h, c1, chat)
        pts, _ = lo.face_points(n, f)
        np.add.at(corr, (pts[:, 0], pts[:, 1], pts[:, 2]), c1 * s * t)
    pts = np.argwhere(np.ones(n, bool))
    lap_thin = _laplacian_update(thin, c, R, pts + R)
    lap_wide = _laplacian_update(wide, c, R, pts + np.asarray(o) + R)
    got = lap_thin + corr[pts[:, 0], pts[:, 1], pts[:, 2]]
    assert np.allclose(got, lap_wide, rtol=1e-12, atol=1e-12)


def test_inputs_on_a_linear_ramp():
    """Along the inward normal p = h + 2 + 0.5*m (m the inward offset index):
    s and the normalised inputs follow in closed form."""
    R = lo.R
    h = 0.25
    n = (12, 12, 12)
    N = n
    idx = np.arange(-R, n[0] + R)
    p = np.broadcast_to((h + 2.0 + 0.5 * idx)[:, None, None], (n[0] + 2 * R, n[1] + 2 * R, n[2] + 2 * R)).copy()
    vp = np.full(N, 0.04)
    X, s, d, v, t = lo.layer_inputs(p, p, vp, (0, 0, 0), n, 0, h, 4.0, np.array([0, 1, 0.5, 0.25, 0.125]))
    for q in range(0, len(d), 37):
        dd = int(d[q])
        i = dd  # low-x face: depth = i
        vals = np.array([2.0 + 0.5 * (i + m - 4) for m in range(9)])
        vals[: 4 - dd] = 0.0  # beyond the edge
        assert s[q] == pytest.approx(np.max(np.abs(vals)))
        assert np.allclose(X[q, :9], vals / s[q]) and np.allclose(X[q, 9:18], vals / s[q])
        assert X[q, 18] == pytest.approx(np.sqrt(0.04 * 4.0))
        exp_t = sum([0, 1, 0.5, 0.25, 0.125][l] * (2.0 + 0.5 * (i - l)) for l in range(dd + 1, 5)) / s[q]
        assert t[q] == pytest.approx(exp_t)


def test_output_bound_closed_form_and_clamp():
    """ymax(d) = sum_{l>d} |c_l / c_1| for the radius-4 central weights
    (8/5, -1/5, 8/315, -1/560): at depth 0 (on the edge) all four terms are
    missing, 1 + 1/8 + 1/63 + 1/896, down to 1/896 at depth 3; corrections()
    clamps y there (a network whose output is a large constant gives exactly
    v c_1 s ymax)."""
    c = np.array([0.0, 8 / 5, -1 / 5, 8 / 315, -1 / 560], np.float32)
    chat = (c / c[1]).astype(np.float32)
    exact = [1 + 1 / 8 + 1 / 63 + 1 / 896, 1 / 8 + 1 / 63 + 1 / 896, 1 / 63 + 1 / 896, 1 / 896]
    for d in range(lo.W):
        assert abs(lo.ymax(chat, d) - exact[d]) <= 1e-7
    n = (20, 18, 22)
    rng = np.random.default_rng(5)
    ext = (n[0] + 2 * lo.R, n[1] + 2 * lo.R, n[2] + 2 * lo.R)
    pn, pm = rng.normal(0, 1, ext), rng.normal(0, 1, ext)
    vp = rng.uniform(0.1, 0.4, n)
    theta = np.zeros((lo.FACES, lo.NP), np.float32)
    theta[:, lo.OFF_B3] = 64.0            # y = 64 everywhere (exact in bf16)
    corr = lo.corrections(theta, pn, pm, vp, n, 0.0, float(c[1]), chat)
    for f in range(lo.FACES):
        X, s, d, v, _ = lo.layer_inputs(pn, pm, vp, (0, 0, 0), n, f, 0.0, float(c[1]))
        bound = np.array([lo.ymax(chat, k) for k in range(lo.W)])[d.astype(int)]
        assert np.allclose(corr[f], v * float(c[1]) * s * bound, rtol=1e-12, atol=0)

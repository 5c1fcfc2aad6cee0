"""Pins the FD oracle (oracle/fd_oracle.c) against the reference's published
outputs and against what the mathematics fixes."""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

from conftest import GOLDEN


def central_weights(L):
    """Exact central second-difference weights of half width L, from the
    moment conditions sum_l w_l l^(2m) = 2 delta_{m,1} (m = 0..L), solved in
    rationals.  Independent of any table."""
    n = L + 1
    A = [[Fraction(0)] * n for _ in range(n)]
    b = [Fraction(0)] * n
    for m in range(n):
        A[m][0] = Fraction(1 if m == 0 else 0)
        for l in range(1, n):
            A[m][l] = Fraction(2 * l ** (2 * m))
        b[m] = Fraction(2 if m == 1 else 0)
    # Gauss-Jordan in rationals
    for c in range(n):
        p = next(r for r in range(c, n) if A[r][c] != 0)
        A[c], A[p] = A[p], A[c]
        b[c], b[p] = b[p], b[c]
        for r in range(n):
            if r != c and A[r][c] != 0:
                f = A[r][c] / A[c][c]
                A[r] = [x - f * y for x, y in zip(A[r], A[c])]
                b[r] -= f * b[c]
    return [b[i] / A[i][i] for i in range(n)]


@pytest.mark.parametrize("L", [1, 2, 3, 4, 5])
def test_fd_coefficients_are_central_weights(oracle, L):
    dx = 10.0
    c = oracle.fd_coefs(L, dx)
    w = central_weights(L)
    for l in range(L + 1):
        exact = float(w[l]) / (dx * dx)
        assert abs(c[l] - exact) <= 2e-7 * abs(exact) + 1e-12, (L, l, c[l], exact)


def test_fd_coefficients_L6_reference_entry(oracle):
    # the reference's sixth weight is 1/1663 (the exact value is -1/16632);
    # the oracle keeps the reference's value, the first five are exact
    c = oracle.fd_coefs(6, 1.0)
    w = central_weights(6)
    for l in range(1, 6):
        assert abs(c[l] - float(w[l])) <= 2e-7 * abs(float(w[l]))
    assert c[6] == np.float32(np.float32(1.0) / np.float32(1663.0))


def test_fd_dt_closed_form(oracle):
    P = oracle.FdProblem(nx=8, model="homogeneous", nsamples=4)
    c = P.cx.astype(np.float64)
    ftmp = 3 * abs(c[0]) + 3 * 2 * np.abs(c[1:]).sum()
    dt = 2 * 0.8 / (np.sqrt(ftmp) * 4500.0)
    assert abs(P.dt - dt) < 1e-6 * dt
    # the published FD runs used dt = 8.05e-4 (test_20240512 logs, header)
    assert abs(P.dt - 8.0508e-4) < 1e-7


def test_ricker_closed_form(oracle):
    L = oracle.lib()
    for order in (0, 1, 2):
        for t in np.linspace(-0.2, 0.4, 301):
            for dexp in (0, 1):
                v = L.orc_ricker(np.float32(t), np.float32(10.0), order, dexp)
                ref = L.orc_ricker_f64(float(np.float32(t)), 10.0, order)
                scale = (2 * (10 * np.pi) ** 2) ** (order / 2.0) if order else 1.0
                assert abs(v - ref) <= 2e-5 * scale + 1e-6 * abs(ref), (order, t, v, ref)


def test_ricker_published_samples(oracle):
    # the published SEM logs print sourceTerm of samples 0/100/200 (f0 = 10 Hz,
    # dt = 1e-3, first-derivative pulse, std::exp(float) translation unit)
    d = json.load(open(os.path.join(GOLDEN, "published_traces.json")))
    hdr = d["result_Cypress_sem_SEQUENTIAL_EPYCROME_01.txt"]["header"]
    src = np.zeros(1000, np.float32)
    oracle.lib().orc_source_term(1000, np.float32(0.001), np.float32(10.0), 1, 0, src)
    for it, txt in hdr["source_samples"]:
        assert "%g" % src[it] == ("%g" % float(txt)).replace("-0", "0") or \
            "%g" % src[it] == txt or ("%g" % src[it]).lstrip("-") == txt.lstrip("-"), (it, src[it], txt)


def test_sponge_table_and_face_rule(oracle):
    nx = ny = nz = 50
    W = 20
    s = np.zeros(nx * ny * nz, np.float32)
    oracle.lib().orc_fd_sponge(nx, ny, nz, W, np.float32(-0.00015), s)
    s = s.reshape(nx, ny, nz)
    f = lambda d: np.float32(np.exp(-0.00015 * d * d))
    # low face: d = W - i ; high face: d = W - (n - i)
    assert s[0, 25, 25] == f(W)
    assert s[W - 1, 25, 25] == f(1)
    assert s[nx - W, 25, 25] == f(0) == np.float32(1)
    assert s[nx - 1, 25, 25] == f(W - 1)
    assert s[25, 3, 25] == f(W - 3)
    assert s[25, 25, nz - 2] == f(W - 2)
    # edges and corners stay 1
    assert s[0, 0, 25] == 1 and s[0, 0, 0] == 1 and s[5, 25, 45] == 1
    # interior stays 1
    assert s[W, W, W] == 1 and s[25, 25, 25] == 1


def test_laplacian_of_quadratic_is_six(oracle):
    # 8th-order weights are exact on polynomials up to degree 9, so with
    # vp = 1, pn = x^2+y^2+z^2 (in grid units), pnm1 = 0 and no sponge one
    # inner3D application gives 2 pn + (coef0 pn + lap) = 2 pn + 6 (dx = 1).
    import ctypes
    P = oracle.FdProblem(nx=16, dx=1.0, model="homogeneous", nsamples=2, sponge_width=0)
    L = P.L
    n = 16
    ax = np.arange(-L, n + L, dtype=np.float64)
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    pn = ((X - 7.3) ** 2 + (Y - 8.1) ** 2 + (Z - 6.7) ** 2).astype(np.float32)
    cur = pn.reshape(-1).copy()
    prev = np.zeros_like(cur)
    vp = np.ones(n ** 3, np.float32)
    oracle.lib().orc_fd_inner3d(ctypes.byref(P.g), P.cx, P.cy, P.cz, P.coef0, vp, prev, cur)
    res = prev.reshape(n + 2 * L, n + 2 * L, n + 2 * L)[L:-L, L:-L, L:-L] - 2 * pn[L:-L, L:-L, L:-L]
    assert np.allclose(res, 6.0, atol=2e-3), (res.min(), res.max())


def test_fd_step_matches_fp64_triple_loop(oracle):
    """One FP32 reference-semantics step vs a brute-force numpy FP64 step."""
    import ctypes
    rng = np.random.default_rng(3)
    P = oracle.FdProblem(nx=12, ny=10, nz=14, model="homogeneous", nsamples=2, sponge_width=3)
    L = P.L
    shp = P.ext_shape()
    cur = rng.standard_normal(shp).astype(np.float32)
    prev = rng.standard_normal(shp).astype(np.float32)
    vp = (0.1 + rng.random(12 * 10 * 14)).astype(np.float32)
    c = P.cx.astype(np.float64)
    C = cur.astype(np.float64)
    lap = np.zeros((12, 10, 14))
    I = (slice(L, -L),) * 3
    for l in range(1, L + 1):
        lap += c[l] * (np.roll(C, -l, 0) + np.roll(C, l, 0))[I]
        lap += c[l] * (np.roll(C, -l, 1) + np.roll(C, l, 1))[I]
        lap += c[l] * (np.roll(C, -l, 2) + np.roll(C, l, 2))[I]
    new = 2 * C[I] - prev.astype(np.float64)[I] + vp.reshape(12, 10, 14) * (P.coef0 * C[I] + lap)
    sp = P.sponge.reshape(12, 10, 14).astype(np.float64)
    a = prev.reshape(-1).copy()
    b = cur.reshape(-1).copy()
    oracle.lib().orc_fd_inner3d(ctypes.byref(P.g), P.cx, P.cy, P.cz, P.coef0, vp, a, b)
    oracle.lib().orc_fd_apply_sponge(ctypes.byref(P.g), P.sponge, a, b)
    got = a.reshape(shp)[I]
    assert np.allclose(got, new * sp, rtol=1e-5, atol=1e-5)
    assert np.array_equal(b.reshape(shp)[I], (cur[I] * P.sponge.reshape(12, 10, 14)).astype(np.float32))
    # halo untouched
    assert np.array_equal(a.reshape(shp)[0], prev[0])


def test_fd_oracle_reproduces_published_prints(oracle):
    """The published CPU logs (150^3, radius 4, homogeneous vp = 1500,
    f0 = 15 Hz, source at the centre) print the source point every 50 steps.
    The oracle reproduces the first 9 prints digit for digit.  From step 450 on
    the logs come from a revision with a different boundary treatment."""
    d = json.load(open(os.path.join(GOLDEN, "published_traces.json")))
    rec = d["result_Cypress_fd_SEQUENTIAL_EPYCROME_01.txt"]
    assert rec["header"]["grid"] == [150, 150, 150]
    P = oracle.FdProblem(nx=150, model="homogeneous")
    assert P.nsamples == 1242
    tr, _ = P.run(401)
    for step, txt in zip(rec["steps"], rec["values"]):
        if step > 400:
            break
        assert "%f" % tr[step] == txt, (step, tr[step], txt)

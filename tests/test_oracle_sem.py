"""Pins the SEM oracle (oracle/sem_oracle.c) by closed forms, invariants and a
brute-force dense element operator."""
import ctypes

import numpy as np
import pytest
from numpy.polynomial import legendre as npleg


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_gll_points_weights_closed_form(oracle, p):
    xi, w, D = oracle.gll(p, ref_rounding=False)
    # interior points are the roots of P_p'
    c = np.zeros(p + 1); c[p] = 1
    roots = np.sort(npleg.legroots(npleg.legder(c))) if p > 1 else np.array([])
    assert np.allclose(xi[1:-1], roots, atol=1e-14)
    assert xi[0] == -1 and xi[-1] == 1
    Pp = npleg.legval(xi, c)
    assert np.allclose(w, 2.0 / (p * (p + 1) * Pp ** 2), rtol=1e-14)
    # quadrature exact to degree 2p-1
    for deg in range(2 * p):
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        assert abs(np.dot(w, xi ** deg) - exact) < 1e-13


def test_gll_reference_weight_table(oracle):
    # SEMQkGLBasisFunctions.hpp:80-117 carries the weights to 8 decimals
    tab = {2: [0.33333333, 1.33333333], 3: [0.16666667, 0.83333333],
           4: [0.1, 0.54444444, 0.71111111], 5: [0.06666667, 0.37847496, 0.55485838]}
    for p, half in tab.items():
        w = oracle.gll(p, ref_rounding=True)[1]
        for i, v in enumerate(half):
            assert w[i] == v and w[p - i] == v
    # and the reference's closed-form points (sqrt5, sqrt(3/7), ...)
    assert oracle.gll(3)[0][1] == pytest.approx(-1 / 2.2360679774997897, abs=3e-16)
    assert oracle.gll(4)[0][1] == pytest.approx(-0.6546536707079771, abs=3e-16)


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_derivative_matrix(oracle, p):
    xi, w, D = oracle.gll(p, ref_rounding=False)
    # D[j, i] = l_j'(xi_i): exact on polynomials of degree <= p
    for deg in range(p + 1):
        f = xi ** deg
        df = deg * xi ** (deg - 1) if deg else np.zeros_like(xi)
        assert np.allclose(f @ D, df, atol=1e-12)
    assert np.allclose(D.sum(axis=0), 0, atol=1e-12)


def make_sem(oracle, ex=3, p=2, fp64=False, model=None, lx=120.0, **kw):
    return oracle.SemProblem(ex=ex, lx=lx, order=p, model=model, nsamples=8,
                             src_point=(lx / 2 + 1, lx / 2 + 1, lx / 2 + 1), fp64=fp64, **kw)


@pytest.mark.parametrize("p", [2, 3, 4])
@pytest.mark.parametrize("fp64", [False, True])
def test_stiffness_kills_constants_and_mass_sums_to_volume(oracle, p, fp64):
    rng = np.random.default_rng(p)
    model = (1500 + 1000 * rng.random(27)).astype(np.float32)
    S = make_sem(oracle, ex=3, p=p, fp64=fp64, model=model)
    y, M = S.apply_k(np.ones(S.nN))
    scale = np.abs(S.apply_k(rng.standard_normal(S.nN))[0]).max()
    assert np.abs(y).max() < (1e-5 if not fp64 else 1e-12) * scale
    h = 120.0 / 3
    vol = (h ** 3 / model.astype(np.float64) ** 2).sum()
    assert abs(M.sum() - vol) < (2e-6 if not fp64 else 1e-12) * vol


@pytest.mark.parametrize("p", [2, 3, 4])
def test_stiffness_symmetric_psd(oracle, p):
    S = make_sem(oracle, ex=2, p=p, fp64=True)
    n = S.nN
    K = np.zeros((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[:] = 0; e[j] = 1
        K[:, j] = S.apply_k(e)[0]
    assert np.allclose(K, K.T, atol=1e-10 * np.abs(K).max())
    ev = np.linalg.eigvalsh(K)
    assert ev.min() > -1e-9 * ev.max()
    assert np.sum(ev < 1e-9 * ev.max()) == 1   # only the constants


@pytest.mark.parametrize("p", [2, 3, 4])
@pytest.mark.parametrize("fp64", [False, True])
def test_quadratic_exactness(oracle, p, fp64):
    """For u = x^2 (+ y^2 + z^2) the assembled K u equals -M (Laplacian u) =
    -2 M per axis on interior nodes, to rounding: GLL quadrature integrates the
    products exactly at these degrees on a box."""
    S = make_sem(oracle, ex=4, p=p, fp64=fp64, lx=160.0)
    x, yy, z = S.node_coords()
    interior = np.zeros(S.nN, bool)
    lst = np.zeros(S.nN, np.int32)
    nI = oracle.lib().orc_sem_interior_nodes(ctypes.byref(S.mesh), lst)
    interior[lst[:nI]] = True
    for u, lap in ((x ** 2, 2.0), (x ** 2 + yy ** 2 + z ** 2, 6.0), (x * z, 0.0)):
        y, M = S.apply_k(u)
        M = M.astype(np.float64) * 1500.0 ** 2  # model c = 1500 divides M
        got = -y.astype(np.float64)[interior] / M[interior]
        tol = 1e-9 if fp64 else 2e-3
        assert np.allclose(got, lap, atol=tol * max(abs(lap), 1) + (0 if fp64 else 2e-3)), \
            (u is None, got.min(), got.max(), lap)


@pytest.mark.parametrize("p", [2, 3, 4])
def test_element_operator_matches_dense_brute_force(oracle, p):
    """One element: the reference-semantics FP32 element loop and the FP64
    sum-factorised one against direct quadrature of grad(phi_i).grad(phi_j)."""
    S32 = make_sem(oracle, ex=1, p=p, lx=40.0)
    S64 = make_sem(oracle, ex=1, p=p, lx=40.0, fp64=True)
    q = p + 1
    xi = oracle.gll(p)[0]
    c = 20 * (xi + 1)
    X = np.tile(c, q * q)
    Z = np.tile(np.repeat(c, q), q)
    Y = np.repeat(c, q * q)
    K = np.zeros(q ** 6)
    Ml = np.zeros(q ** 3)
    oracle.lib().orc_sem_dense_element_f64(p, X, Y, Z, K, Ml)
    K = K.reshape(q ** 3, q ** 3)
    rng = np.random.default_rng(7)
    u = rng.standard_normal(q ** 3)
    ref = K @ u
    nodes = S32.nodes[0]
    pu = np.zeros(S32.nN); pu[nodes] = u
    y32 = S32.apply_k(pu.astype(np.float32))[0][nodes]
    y64 = S64.apply_k(pu)[0][nodes]
    assert np.allclose(y64, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    assert np.abs(y32 - ref).max() < 2e-6 * np.abs(ref).max()


def test_dense_brute_force_detects_general_hex(oracle):
    """On a sheared element the dense operator still kills constants and is
    symmetric (the operator is defined for general hexahedra)."""
    p, q = 2, 3
    xi = oracle.gll(p)[0]
    c = 10 * (xi + 1)
    X = np.tile(c, q * q); Z = np.tile(np.repeat(c, q), q); Y = np.repeat(c, q * q)
    X = X + 0.3 * Y  # shear
    K = np.zeros(q ** 6); Ml = np.zeros(q ** 3)
    oracle.lib().orc_sem_dense_element_f64(p, X, Y, Z, K, Ml)
    K = K.reshape(q ** 3, q ** 3)
    assert np.allclose(K @ np.ones(q ** 3), 0, atol=1e-10)
    assert np.allclose(K, K.T, atol=1e-10)
    assert abs(Ml.sum() - 20 ** 3) < 1e-8


def test_fp32_reference_semantics_tracks_fp64(oracle):
    """Reference-semantics FP32 run vs the FP64 oracle on a small problem."""
    kw = dict(ex=6, lx=240.0, order=2, dt=0.002, f0=20.0, nsamples=120,
              src_point=(121.0, 121.0, 121.0))
    a = oracle.SemProblem(**kw)
    b = oracle.SemProblem(fp64=True, **kw)
    ta, _ = a.run(120)
    tb, _ = b.run(120)
    err = np.linalg.norm(a.current() - b.current()) / np.linalg.norm(b.current())
    assert err < 1e-4, err
    assert np.linalg.norm(ta - tb) / np.linalg.norm(tb) < 1e-4

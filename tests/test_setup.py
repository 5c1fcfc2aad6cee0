"""Host-side setup of libphc.so (coefficients, dt, wavelet, models) equals the
oracle's restatement bit for bit (no device work)."""
import numpy as np
import pytest


@pytest.mark.parametrize("L", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("dx", [10.0, 7.0, 1.0])
def test_coefficients_dt_coef0(phc, oracle, L, dx):
    a = phc.fd_coefs(L, dx)
    b = oracle.fd_coefs(L, dx)
    assert np.array_equal(a, b)
    c0a = phc.lib().phc_fd_coef0(a, L, a, L, a, L)
    c0b = oracle.lib().orc_fd_coef0(b, L, b, L, b, L)
    assert c0a == c0b
    for cfl in (0.8, 0.4):
        assert phc.lib().phc_fd_compute_dt(4500.0, a, L, a, L, a, L, cfl) == \
            oracle.lib().orc_fd_compute_dt(4500.0, b, L, b, L, b, L, cfl)


@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("variant", [0, 1])
def test_source_term(phc, oracle, order, variant):
    a = phc.source_term(2000, np.float32(4.0e-4), 10.0, order, variant)
    b = np.zeros(2000, np.float32)
    # oracle flag is use_double_exp (1 = double), product flag PHC_EXP_DOUBLE = 0
    oracle.lib().orc_source_term(2000, np.float32(4.0e-4), 10.0, order, 1 - variant, b)
    assert np.array_equal(a, b)


def test_two_layer_model_matches_oracle(phc, oracle):
    s = phc.FdSetup(nx=20, ny=18, nz=22, model="two_layer", nsamples=4)
    P = oracle.FdProblem(nx=20, ny=18, nz=22, model="layered2", nsamples=4)
    assert np.array_equal(s.vp, P.vp)
    assert np.array_equal(s.rhs, P.rhs)
    assert s.dt == P.dt


def test_layered_model(phc):
    s = phc.FdSetup(nx=8, ny=8, nz=512, model=("layered", 6, 12345), nsamples=4, cfl=0.4)
    itf = list(s.interfaces)
    assert itf == sorted(itf) and len(set(itf)) == 5 and 0 < itf[0] and itf[-1] < 512
    v = np.sqrt(s.vp.reshape(8, 8, 512)[0, 0] / s.dt ** 2)
    assert abs(v[0] - 1500) < 0.5 and abs(v[-1] - 4500) < 1.5
    assert np.all(np.diff(v) >= -1e-3)
    assert len(np.unique(np.round(v))) == 6

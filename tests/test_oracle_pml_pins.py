"""The oracle's PML step pinned point by point (CPU only).

The reference has no published PML run, so its pml3D (FDTDkernels.hpp:36-73)
and init_eta (FDTDutils.hpp:78-190) are restated here a second time, in
float64 numpy, straight from the reference's expressions: the 3D eta volume
as init_eta builds it (pml_profile_init ramps, then pml_profile_extend_all
over the six regions), and one pml3D / inner3D step over the whole grid.  One
oracle step from a non-trivial state (smooth fields with gradients along all
three axes, phi != 0, eta up to its full scale) must agree with it, and
plausible mistakes -- hd2 doubled, a dropped /(1+eta) or /(1+2 eta), eta read
one point off, a sign flipped -- must not."""
import numpy as np
import pytest

TOL = 1e-5


def eta_volume(nx, ny, nz, nd, scale):
    """FDTDutils.hpp:129-190 init_eta: the (nx+2)(ny+2)(nz+2) volume."""
    def profile(n):
        # pml_profile_init(profile, 0, n+1, nd, nd, scale), FDTDutils.hpp:78-104
        i_min, i_max = 0, n + 1
        nn = i_max - i_min + 1
        shift = i_min - 1
        first_beg, first_end = 1 + shift, nd + shift
        last_beg, last_end = nn - nd + 1 + shift, nn + shift
        p = np.zeros(n + 2, np.float64)
        tmp = np.float32(scale) / np.float32((first_end - first_beg + 1) ** 2)
        for i in range(1, first_end - first_beg + 2):
            p[first_end - i + 1] = np.float32(i * i) * tmp
        for i in range(1, last_end - last_beg + 2):
            p[last_beg + i - 1] = np.float32(i * i) * tmp
        return p
    ex, ey, ez = profile(nx), profile(ny), profile(nz)
    E = np.zeros((nx + 2, ny + 2, nz + 2), np.float64)

    def extend(xb, xe, yb, ye, zb, ze):
        # pml_profile_extend, FDTDutils.hpp:108-121 (one ghost layer)
        for ix in range(xb - 1, xe + 2):
            for iy in range(yb - 1, ye + 2):
                E[ix, iy, zb - 1:ze + 2] = ex[ix] + ey[iy] + ez[zb - 1:ze + 2]
    # region bounds of FDTDinit.hpp:174-198, shifted by one as init_eta passes them
    x1, x2, x5, x6 = 0, nd, nx - nd, nx
    y1, y2, y3, y4, y5, y6 = 0, nd, nd, ny - nd, ny - nd, ny
    z1, z2, z3, z4, z5, z6 = 0, nd, nd, nz - nd, nz - nd, nz
    xmin, xmax, ymin, ymax = 1, nx, 1, ny
    extend(xmin, xmax, ymin, ymax, z1 + 1, z2)          # top
    extend(xmin, xmax, ymin, ymax, z5 + 1, z6)          # bottom
    extend(xmin, xmax, y1 + 1, y2, z3 + 1, z4)          # front
    extend(xmin, xmax, y5 + 1, y6, z3 + 1, z4)          # back
    extend(x1 + 1, x2, y3 + 1, y4, z3 + 1, z4)          # left
    extend(x5 + 1, x6, y3 + 1, y4, z3 + 1, z4)          # right
    return E, (ex, ey, ez)


def reference_step(P, E, nd, hd2, mutate=None):
    """One computeOneStepPML (FDTDkernels.hpp:86-104) in float64: pml3D on the
    slabs, inner3D inside.  Returns (new level, new phi) on the interior."""
    nx, ny, nz = P.nx, P.ny, P.nz
    lx, ly, lz = P.lh
    cb = P.p1.reshape(P.ext_shape()).astype(np.float64)
    ca = P.p0.reshape(P.ext_shape()).astype(np.float64)
    vp = P.vp.reshape(nx, ny, nz).astype(np.float64)
    phi = P.phi.reshape(nx, ny, nz).astype(np.float64)
    I = np.s_[lx:lx + nx, ly:ly + ny, lz:lz + nz]

    def sh(a, di, dj, dk):
        return a[lx + di:lx + di + nx, ly + dj:ly + dj + ny, lz + dk:lz + dk + nz]
    lap = P.coef0 * cb[I]
    for l in range(1, lx + 1):
        lap = lap + float(P.cx[l]) * (sh(cb, l, 0, 0) + sh(cb, -l, 0, 0))
    for l in range(1, ly + 1):
        lap = lap + float(P.cy[l]) * (sh(cb, 0, l, 0) + sh(cb, 0, -l, 0))
    for l in range(1, lz + 1):
        lap = lap + float(P.cz[l]) * (sh(cb, 0, 0, l) + sh(cb, 0, 0, -l))
    e = E[1:nx + 1, 1:ny + 1, 1:nz + 1]
    if mutate == "eta_shift":
        e = E[2:nx + 2, 1:ny + 1, 1:nz + 1]
    pc, pm = cb[I], ca[I]
    e2 = -e * e if mutate == "sign" else e * e
    num = (2.0 - e2 + 2.0 * e) * pc - pm + vp * (phi + lap)
    pnew = num if mutate == "no_p_div" else num / (1.0 + 2.0 * e)
    h = np.array(hd2, np.float64) * (2.0 if mutate == "hd2x2" else 1.0)
    tg = ((E[2:nx + 2, 1:ny + 1, 1:nz + 1] - E[0:nx, 1:ny + 1, 1:nz + 1]) * (sh(cb, 1, 0, 0) - sh(cb, -1, 0, 0)) * h[0]
          + (E[1:nx + 1, 2:ny + 2, 1:nz + 1] - E[1:nx + 1, 0:ny, 1:nz + 1]) * (sh(cb, 0, 1, 0) - sh(cb, 0, -1, 0)) * h[1]
          + (E[1:nx + 1, 1:ny + 1, 2:nz + 2] - E[1:nx + 1, 1:ny + 1, 0:nz]) * (sh(cb, 0, 0, 1) - sh(cb, 0, 0, -1)) * h[2])
    phin = (phi - tg) if mutate == "no_phi_div" else (phi - tg) / (1.0 + e)
    inner = np.zeros((nx, ny, nz), bool)
    inner[nd:nx - nd, nd:ny - nd, nd:nz - nd] = True
    pin = 2.0 * pc - pm + vp * lap
    return np.where(inner, pin, pnew), np.where(inner, phi, phin), inner


@pytest.mark.parametrize("lh", [(4, 4, 4), (3, 2, 4)])
def test_pml_step_matches_reference_expressions(oracle, lh):
    nx, ny, nz, nd = 22, 19, 24, 5
    lx, ly, lz = lh
    P = oracle.FdProblem(nx=nx, ny=ny, nz=nz, L=lx, ly=ly, lz=lz, f0=30.0, model="layered2", nsamples=4,
                         sponge_width=0, pml=True, ndamp=nd)
    P.rhs[:] = 0.0
    rng = np.random.default_rng(7)
    X, Y, Z = np.meshgrid(np.arange(nx + 2 * lx), np.arange(ny + 2 * ly), np.arange(nz + 2 * lz), indexing="ij")
    cb = np.sin(0.31 * X + 0.2) * np.cos(0.23 * Y - 0.4) + 0.5 * np.sin(0.17 * Z + 0.05 * X * Y / 10.0)
    ca = 0.9 * cb + 0.05 * rng.standard_normal(cb.shape)
    P.p1[:] = cb.astype(np.float32).reshape(-1)
    P.p0[:] = ca.astype(np.float32).reshape(-1)
    phi0 = (1e-5 * rng.standard_normal(nx * ny * nz)).astype(np.float32)
    P.phi[:] = phi0
    scale = float(P.eta[0].max())
    E, profs = eta_volume(nx, ny, nz, nd, scale)
    # the oracle's eta = sum of the three profiles equals the volume init_eta builds
    assert np.array_equal(profs[0].astype(np.float32), P.eta[0])
    S = (P.eta[0][:, None, None].astype(np.float64) + P.eta[1][None, :, None] + P.eta[2][None, None, :])
    assert np.allclose(S, E, rtol=1e-6, atol=1e-7)
    assert E.max() > 0.3  # the ramps reach their full scale
    expect_p, expect_phi, inner = reference_step(P, E, nd, P.hd2)
    P.run(1)
    got_p = P.current()[lx:lx + nx, ly:ly + ny, lz:lz + nz].astype(np.float64)
    got_phi = P.phi.reshape(nx, ny, nz).astype(np.float64)
    sp, sf = np.abs(expect_p).max(), np.abs(expect_phi[~inner]).max()
    assert sf > 1e-7
    assert np.abs(got_p - expect_p).max() <= TOL * sp
    assert np.abs(got_phi - expect_phi).max() <= TOL * sf
    # each planted mistake moves the slab values far outside that tolerance
    for m in ("hd2x2", "no_phi_div", "no_p_div", "eta_shift", "sign"):
        Q = oracle.FdProblem(nx=nx, ny=ny, nz=nz, L=lx, ly=ly, lz=lz, f0=30.0, model="layered2", nsamples=4,
                             sponge_width=0, pml=True, ndamp=nd)
        Q.p1[:] = cb.astype(np.float32).reshape(-1)
        Q.p0[:] = ca.astype(np.float32).reshape(-1)
        Q.phi[:] = phi0
        mp, mf, _ = reference_step(Q, E, nd, Q.hd2, mutate=m)
        dev = max(np.abs(got_p - mp).max() / sp, np.abs(got_phi - mf).max() / sf)
        assert dev > 100 * TOL, (m, dev)

"""FP64 FD loop: the GPU path (phc_fd64_*) against the CPU FP64 restatement
(oracle/fd_oracle.c orc_fd_run_f64), and that restatement pinned by the
reference-rounding FP32 oracle and by a stationary constant field."""
import ctypes

import numpy as np
import pytest

from oracle import FdGrid


def sponge64(n, W, alpha):
    """Face-slab sponge (FDTDinit.hpp:343-395) with the factor in double."""
    def axis(m):
        f = np.ones(m)
        for i in range(m):
            d = -1
            if i >= m - W:
                d = W - (m - i)
            elif i < W:
                d = W - i
            if d >= 0:
                f[i] = np.exp(np.float64(np.float32(alpha)) * float(d) ** 2)
        return f
    nx, ny, nz = n
    fx, fy, fz = axis(nx), axis(ny), axis(nz)
    inI = (np.arange(nx) >= W) & (np.arange(nx) < nx - W)
    inJ = (np.arange(ny) >= W) & (np.arange(ny) < ny - W)
    inK = (np.arange(nz) >= W) & (np.arange(nz) < nz - W)
    I, J, K = np.meshgrid(inI, inJ, inK, indexing="ij")
    s = np.ones(n)
    s = np.where(J & K & ~I, fx[:, None, None], s)
    s = np.where(K & I & ~J, fy[None, :, None], s)
    s = np.where(J & I & ~K, fz[None, None, :], s)
    return s


def run_oracle64(oracle, s, nsteps, rec=None, init=1e-6):
    L = 4
    g = FdGrid()
    g.nx, g.ny, g.nz = s.nx, s.ny, s.nz
    g.lx = g.ly = g.lz = L
    g.xs, g.ys, g.zs = s.xs, s.ys, s.zs
    g.x3, g.x4, g.y3, g.y4, g.z3, g.z4 = 0, s.nx, 0, s.ny, 0, s.nz
    ext = (s.nx + 2 * L) * (s.ny + 2 * L) * (s.nz + 2 * L)
    v0 = float(np.float32(init))
    p0 = np.full(ext, v0)
    p1 = np.full(ext, v0)
    cur = ctypes.c_int(1)
    tr = np.zeros(nsteps)
    nrec = 0 if rec is None else len(rec)
    rv = np.zeros(max(nrec, 1) * nsteps)
    rij = None if rec is None else np.ascontiguousarray(rec, np.int32).reshape(-1)
    sp = sponge64((s.nx, s.ny, s.nz), s.sponge_width, s.sponge_alpha).reshape(-1)
    oracle.lib().orc_fd_run_f64(ctypes.byref(g), s.cx.astype(np.float64), s.cy.astype(np.float64),
                                s.cz.astype(np.float64), float(s.coef0), s.vp.astype(np.float64), sp,
                                s.rhs.astype(np.float64), 0, nsteps, p0, p1, ctypes.byref(cur),
                                tr.ctypes.data_as(ctypes.c_void_p), nrec,
                                None if rij is None else rij.ctypes.data_as(ctypes.c_void_p),
                                rv.ctypes.data_as(ctypes.c_void_p))
    shp = (s.nx + 2 * L, s.ny + 2 * L, s.nz + 2 * L)
    curf = (p1 if cur.value else p0).reshape(shp)
    prev = (p0 if cur.value else p1).reshape(shp)
    return tr, (rv.reshape(nrec, nsteps) if nrec else None), curf, prev


def test_oracle64_tracks_the_fp32_oracle(phc, oracle):
    s = phc.FdSetup(30, 26, 34, L=4, f0=15.0, sponge_width=6, model=("layered", 3, 5), nsamples=60)
    tr64, _, c64, _ = run_oracle64(oracle, s, 50)
    P = oracle.FdProblem(nx=30, ny=26, nz=34, L=4, xs=s.xs, ys=s.ys, zs=s.zs, f0=15.0, model=s.vp.copy(),
                         nsamples=60, sponge_width=6)
    tr32, _ = P.run(50)
    assert np.linalg.norm(tr64 - tr32) <= 1e-5 * np.linalg.norm(tr64)
    c32 = P.current().astype(np.float64)
    assert np.linalg.norm(c64 - c32) <= 1e-5 * np.linalg.norm(c64 - 1e-6)


def test_oracle64_constant_field_is_stationary(phc, oracle):
    """With coef0 = -2 sum(c) exactly (the reference forms it from float sums,
    FDTDinit.hpp:223-227, which leaves a ~1e-9 residual), a constant field
    with no source and no sponge stays constant to rounding."""
    s = phc.FdSetup(16, 17, 18, L=4, sponge_width=0, model="homogeneous", nsamples=10)
    s.rhs[:] = 0.0
    s.coef0 = -2.0 * float(np.sum(s.cx[1:].astype(np.float64)) + np.sum(s.cy[1:].astype(np.float64)) +
                           np.sum(s.cz[1:].astype(np.float64)))
    tr, _, c, p = run_oracle64(oracle, s, 8, init=0.5)
    assert np.allclose(c, 0.5, rtol=0, atol=1e-14) and np.allclose(tr, 0.5, atol=1e-14)


@pytest.mark.gpu
def test_gpu_fp64_matches_oracle64(gpu, oracle):
    s = gpu.FdSetup(37, 29, 70, L=4, xs=12, ys=20, zs=33, f0=15.0, sponge_width=8, model=("layered", 3, 7),
                    nsamples=60)
    rec = np.array([[12, 20, 33], [0, 0, 0], [36, 28, 69], [20, 10, 5]], np.int32)
    sol = gpu.Fd64Solver(s)
    sol.set_receivers(rec)
    src, rv = sol.run(45)
    tr, orv, oc, op = run_oracle64(oracle, s, 45, rec)
    assert np.allclose(src, tr, rtol=1e-13, atol=1e-300)
    assert np.allclose(rv, orv, rtol=1e-12, atol=1e-18)
    assert np.max(np.abs(sol.field(True) - oc)) <= 1e-13 * np.max(np.abs(oc))
    assert np.max(np.abs(sol.field(False) - op)) <= 1e-13 * np.max(np.abs(op))
    # continuation
    src2, _ = sol.run(10)
    tr2, _, oc2, _ = run_oracle64(oracle, s, 55, rec)
    assert np.allclose(src2, tr2[45:], rtol=1e-13, atol=1e-300)

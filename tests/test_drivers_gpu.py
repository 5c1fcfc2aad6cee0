"""The C++ drop-in drivers (fd_hip, sem_hip: the reference's classes and argv
on libphc.so) end to end: their prints against the reference's published
prints and against the oracle."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu
PKG = os.path.join(ROOT, "proxy-geos-hc_amd")


def run(exe, args, cwd):
    r = subprocess.run([os.path.join(PKG, exe)] + args, cwd=cwd, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    return r.stdout


@pytest.mark.parametrize("fused", [False, True])
def test_fd_driver_prints_published_values(gpu, tmp_path, fused):
    """fd_hip -nx 150 -homogeneous reproduces the published source-point prints
    of the reference's CPU builds through step 400 (the later prints come from
    a revision with another boundary treatment)."""
    print("\n[drivers] fd_hip fused=%d" % fused, flush=True)
    args = ["-nx", "150", "-homogeneous", "-nsteps", "401"] + (["-fused"] if fused else [])
    out = run("fd_hip", args, str(tmp_path))
    got = dict((int(m.group(1)), m.group(2)) for m in
               re.finditer(r"TimeStep=(\d+)\t; Pressure value at source \[75 75 75\] = (\S+)", out))
    d = json.load(open(os.path.join(GOLDEN, "published_traces.json")))
    rec = d["result_Cypress_fd_SEQUENTIAL_EPYCROME_01.txt"]
    n = 0
    for step, txt in zip(rec["steps"], rec["values"]):
        if step > 400:
            break
        assert got[step] == txt, (step, got[step], txt)
        n += 1
    assert n == 9
    assert "Elapsed Kernel Time" in out


def test_fd_driver_snapshots(gpu, oracle, tmp_path):
    out = run("fd_hip", ["-nx", "40", "-xs", "20", "-savesnapshots", "-nsteps", "51", "-sponge", "8"],
              str(tmp_path))
    snap = np.fromfile(str(tmp_path / "snapshot_it_50.H@"), np.float32)
    assert snap.size == 40 * 40
    P = oracle.FdProblem(nx=40, xs=20, sponge_width=8)
    P.run(51)
    cur = P.current()[4:-4, 4:-4, 4:-4]          # (i, j, k) interior
    ref = cur[:, 20, :].T.reshape(-1)            # k slowest, i fastest
    assert np.array_equal(snap, ref)


@pytest.mark.parametrize("fused", [False, True])
def test_fd_driver_pml(gpu, oracle, tmp_path, fused):
    """fd_hip -usePML (FDTDmain.cpp:64-71 computeOneStepPML): the prints and
    the last snapshot equal the PML oracle's values bit for bit."""
    print("\n[drivers] fd_hip -usePML fused=%d" % fused, flush=True)
    args = ["-nx", "48", "-xs", "21", "-f0", "30", "-usePML", "-nsteps", "151", "-savesnapshots"]
    out = run("fd_hip", args + (["-fused"] if fused else []), str(tmp_path))
    assert "usePML=1" in out
    got = dict((int(m.group(1)), float(m.group(2))) for m in
               re.finditer(r"TimeStep=(\d+)\t; Pressure value at source \[21 21 21\] = (\S+)", out))
    P = oracle.FdProblem(nx=48, xs=21, ys=21, zs=21, f0=30.0, sponge_width=0, pml=True)
    tr, _ = P.run(151)
    for step in (0, 50, 100, 150):
        assert got[step] == float("%f" % tr[step]), step
    snap = np.fromfile(str(tmp_path / "snapshot_it_150.H@"), np.float32)
    ref = P.current()[4:-4, 4:-4, 4:-4][:, 24, :].T.reshape(-1)
    assert np.array_equal(snap, ref)


def test_sem_driver_prints(gpu, oracle, tmp_path):
    """sem_hip 8 320 -order 2 ... prints the source-node value every 100 steps;
    compared with the FP32 reference-semantics oracle (%g, 6 digits)."""
    print("\n[drivers] sem_hip", flush=True)
    out = run("sem_hip", ["10", "2000", "-order", "2", "-dt", "0.001", "-f0", "20", "-nsteps", "301",
                          "-receivers", "-fused"], str(tmp_path))
    got = {int(m.group(1)): float(m.group(2)) for m in
           re.finditer(r"TimeStep=(\d+);  pnGlobal @ elementSource location \d+ after computeOneStep = (\S+)", out)}
    O = oracle.SemProblem(ex=10, lx=2000.0, order=2, dt=0.001, f0=20.0, nsamples=301,
                          src_point=(1001.0, 1001.0, 1001.0))
    tr, _ = O.run(301)
    for step in (0, 100, 200, 300):
        assert abs(got[step] - tr[step]) <= 1e-4 * np.abs(tr).max() + 1e-5 * abs(tr[step]), (step, got[step], tr[step])
    lines = open(str(tmp_path / "receivers.txt")).read().split("\n")
    assert len([l for l in lines if l.strip()]) == 301


@pytest.mark.parametrize("fused", [False, True])
def test_sem_driver_snapshots(gpu, oracle, tmp_path, fused):
    """sem_hip -snapshots writes the reference's ASCII snapshot (SEMmesh.cpp:
    624-654): plane iy = ny/2 - 1, lines "ix*hx iz*hz value" with hx the int
    element size, x-major, stream formatting."""
    print("\n[drivers] sem_hip -snapshots fused=%d" % fused, flush=True)
    args = ["6", "1000", "-order", "3", "-dt", "0.0008", "-f0", "25", "-nsteps", "201", "-snapshots"]
    run("sem_hip", args + (["-fused"] if fused else []), str(tmp_path))
    O = oracle.SemProblem(ex=6, lx=1000.0, order=3, dt=0.0008, f0=25.0, nsamples=201,
                          src_point=(1001.0, 1001.0, 1001.0))
    O.run(201)
    nx = ny = nz = 6 * 3 + 1
    field = O.current()           # the newest level (level i1 when the reference prints)
    lines = [l.split() for l in open(str(tmp_path / "snapshot200")).read().split("\n") if l.strip()]
    assert len(lines) == nx * nz
    hx = int(1000.0 / 6)          # 166: SEMmesh::getDx returns the element size as an int
    off = nx * nz * (ny // 2 - 1)
    scale = np.abs(field).max()
    for n, (x, z, v) in enumerate(lines):
        i, j = n // nz, n % nz
        assert x == "%g" % (i * hx) and z == "%g" % (j * hx)
        assert abs(float(v) - field[off + i + j * nx]) <= 1e-4 * scale + 1e-5 * abs(field[off + i + j * nx])
    assert os.path.exists(str(tmp_path / "snapshot0")) and os.path.exists(str(tmp_path / "snapshot100"))


@pytest.mark.parametrize("fused", [False, True])
def test_sem_driver_2d(gpu, oracle, tmp_path, fused):
    """sem_hip 65 2000 -2d: the reference's SEM2D build on its published mesh
    (65 x 65 elements of order 3); header lines as the reference prints them,
    source element from SEMmesh::getElementNumberFromPoints with the float
    element size of the checked-in revision, prints and the ASCII snapshot of
    the single node plane against the SEM2D oracle."""
    print("\n[drivers] sem_hip -2d fused=%d" % fused, flush=True)
    args = ["65", "2000", "-2d", "-order", "3", "-dt", "0.001", "-f0", "10", "-nsteps", "301", "-snapshots"]
    out = run("sem_hip", args + (["-fused"] if fused else []), str(tmp_path))
    assert "ex, ey, ez=65, 0, 65" in out and "nx, ny, nz=196, 1, 196" in out
    m = re.search(r"Element number for the source location: (\d+)", out)
    e = int(m.group(1))
    assert e == 32 + 32 * 65  # 1001 / (2000/65) = 32.5 on both axes
    got = {int(m.group(1)): float(m.group(2)) for m in
           re.finditer(r"TimeStep=(\d+);  pnGlobal @ elementSource location \d+ after computeOneStep = (\S+)", out)}
    O = oracle.Sem2dProblem(ex=65, order=3, lx=2000.0, dt=1e-3, f0=10.0, nsamples=301, src_element=e)
    tr, _ = O.run(301)
    for step in (0, 100, 200, 300):
        assert abs(got[step] - tr[step]) <= 1e-4 * np.abs(tr).max() + 1e-5 * abs(tr[step]), (step, got[step], tr[step])
    field = O.current()
    lines = [l.split() for l in open(str(tmp_path / "snapshot300")).read().split("\n") if l.strip()]
    assert len(lines) == 196 * 196
    scale = np.abs(field).max()
    for n in range(0, len(lines), 97):
        x, z, v = lines[n]
        i, j = n // 196, n % 196
        assert x == "%g" % (i * 30) and z == "%g" % (j * 30)
        assert abs(float(v) - field[i + j * 196]) <= 1e-4 * scale + 1e-5 * abs(field[i + j * 196])

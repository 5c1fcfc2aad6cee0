"""Sanitizer runs on the CPU (no GPU):

* the oracle under AddressSanitizer + UndefinedBehaviorSanitizer, and under
  ThreadSanitizer with the OpenMP runtime's race annotations: the FD loops,
  the PML step, the SEM setup and the colour-parallel SEM element loop on 4
  threads.  The same build with a race planted in that element loop (the
  reference's unsynchronised OpenMP scatter, SEMmacros.hpp:50-56) must be
  reported, so the clean run means something;
* libphc.so's host code under AddressSanitizer: setup routines, the create
  calls' validation paths, schedule tables, colour lists, the drop-in argv
  parser, mesh helpers and the device-bound array's host rules
  (tests/cpp/host_asan_check.cpp)."""
import os
import shutil
import subprocess
import sys

import pytest

from conftest import ROOT

LLVM = "/opt/rocm/llvm"


def make(target):
    return subprocess.run(["make", "-s", "-f", os.path.join(ROOT, "oracle", "sanitize.mk"), target],
                          stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)


def test_oracle_asan_ubsan():
    if not shutil.which("gcc"):
        pytest.skip("gcc not found")
    r = make("asan")
    assert r.returncode == 0 and "SANITIZE OK" in r.stdout, r.stdout[-3000:]


def test_oracle_tsan_and_planted_race():
    if not os.path.exists(os.path.join(LLVM, "lib", "libarcher.so")):
        pytest.skip("LLVM OpenMP race annotations (libarcher) not found")
    r = make("tsan")
    assert r.returncode == 0 and "SANITIZE OK" in r.stdout, r.stdout[-3000:]
    r = make("tsan-planted")
    assert r.returncode == 0, r.stdout[-3000:]
    assert "ThreadSanitizer: data race" in r.stdout and "assemble_colours" in r.stdout


def test_library_host_code_asan():
    sys.path.insert(0, os.path.join(ROOT, "proxy-geos-hc_amd"))
    import build
    exe = build.sanitize_host()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300, env=env)
    assert r.returncode == 0 and "HOST ASAN OK" in r.stdout, r.stdout[-3000:]

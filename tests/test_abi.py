"""The C ABI library loads and exports every function include/*.h declares
(no device work: runs without a GPU)."""
import glob
import os
import re

from conftest import ROOT


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(phc_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported(phc):
    L = phc.lib()
    missing = [n for n in sorted(declared_functions()) if not hasattr(L, n)]
    assert not missing, missing
    assert len(declared_functions()) > 20


def test_every_declared_symbol_has_a_binding(phc):
    missing = [n for n in sorted(declared_functions()) if n not in phc.SIGNATURES]
    assert not missing, missing


def test_errors_are_reported_without_a_device(phc):
    import ctypes
    p = phc.FdParams()
    h = ctypes.c_void_p()
    st = phc.lib().phc_fd_create(ctypes.byref(p), ctypes.byref(h))
    assert st == -1
    assert b"empty" in phc.lib().phc_last_error()
    assert phc.lib().phc_fd_run(None, 0, 1, None) == -1


def test_sem2d_parameters_checked_without_a_device(phc):
    import ctypes
    p = phc.SemParams()
    p.ex, p.ey, p.ez, p.order = 4, 4, 4, 2
    h = ctypes.c_void_p()
    assert phc.lib().phc_sem2d_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert b"ey = 0" in phc.lib().phc_last_error()
    p.ey = 0
    p.ez = 1
    assert phc.lib().phc_sem2d_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert b"2 x 2" in phc.lib().phc_last_error()

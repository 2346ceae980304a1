"""The C-ABI library loads and exports every entry point declared in include/fdga.h
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import fdga
from fdga import _lib as L


def declared_functions():
    text = open(L.HEADER).read()
    return sorted(set(re.findall(r"\b(fdga_\w+)\s*\(", text)))


def test_library_built():
    assert os.path.exists(L.LIB_PATH), "run __graft_entry__.build()"


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_functions()) == set(L.SIGNATURES)


def test_no_device_fails_loudly():
    """Without a HIP device the context cannot be created and the error says so."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r); import fdga\n"
            "try:\n"
            "    fdga.Solver(nG=4, nK1=3, nK2=(2, 2), nK3=(1, 1), T=0.1, U=1.0, mPinu=1)\n"
            "    print('DEVICE')\n"
            "except fdga.FdgaError as e:\n"
            "    print('ERR', e)\n") % os.path.dirname(os.path.dirname(L.__file__))
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=120)
    assert out.stdout.startswith("ERR") or out.stdout.startswith("DEVICE"), out.stdout + out.stderr

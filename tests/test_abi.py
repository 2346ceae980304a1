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


def _header_dims():
    """(name, C type) of fdga_dims in declaration order, from the header text."""
    text = open(L.HEADER).read()
    body = re.search(r"typedef struct fdga_dims \{(.*?)\} fdga_dims;", text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        ctype, names = decl.split(None, 1)
        fields += [(n.strip(), ctype) for n in names.split(",")]
    return fields


def test_dims_mirrors_match_the_header():
    """The Python ctypes and the Julia mirrors of fdga_dims have the header's fields, order and
    types (a shifted or retyped field would hand the library a corrupted box)."""
    hdr = _header_dims()
    assert len(hdr) == 22
    py = [(n.rstrip("_"), t) for n, t in L.Dims._fields_]
    ctypes_of = {"int32_t": ctypes.c_int32, "double": ctypes.c_double}
    assert py == [(n, ctypes_of[t]) for n, t in hdr]
    off = 0   # natural C layout: each field at a multiple of its size, the struct a multiple of 8
    for _, t in hdr:
        sz = 4 if t == "int32_t" else 8
        off = -(-off // sz) * sz + sz
    assert ctypes.sizeof(L.Dims) == -(-off // 8) * 8
    jl_path = os.path.join(os.path.dirname(L.HEADER), "..", "fddgasolver.jl_amd", "julia", "FdgaHIP.jl")
    jl = open(jl_path).read()
    body = re.search(r"struct FdgaDims(.*?)\nend", jl, re.S).group(1)
    jl_fields = re.findall(r"(\w+)::(Int32|Float64)", body)
    julia_of = {"int32_t": "Int32", "double": "Float64"}
    assert jl_fields == [(n, julia_of[t]) for n, t in hdr]

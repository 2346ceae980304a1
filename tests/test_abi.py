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


def _c_class(decl):
    decl = decl.strip()
    if decl.startswith("const char*"):
        return "str"
    if "*" in decl:
        return "ptr"
    return {"int": "i32", "int32_t": "i32", "int64_t": "i64", "double": "f64"}[decl.split()[0]]


def _header_signatures():
    """name -> (return class, [argument classes]) from the header's declarations."""
    text = re.sub(r"/\*.*?\*/", "", open(L.HEADER).read(), flags=re.S)
    sigs = {}
    for ret, name, params in re.findall(r"^\s*(const char\*|int64_t|int|double)\s+(fdga_\w+)\s*\(([^)]*)\);", text,
                                        re.M):
        sigs[name] = (_c_class(ret), [_c_class(p) for p in params.split(",") if p.strip()])
    return sigs


def test_python_signatures_match_the_header():
    """argtypes/restype of every ctypes binding agree with the header's declaration."""
    def cls(t):
        if t in (ctypes.c_int, ctypes.c_int32):
            return "i32"
        if t in (ctypes.c_int64, ctypes.c_longlong):
            return "i64"
        if t is ctypes.c_double:
            return "f64"
        if t is ctypes.c_char_p:
            return "str"
        assert t is ctypes.c_void_p or issubclass(t, ctypes._Pointer), t
        return "ptr"
    hdr = _header_signatures()
    assert set(hdr) == set(L.SIGNATURES)
    for name, (res, args) in L.SIGNATURES.items():
        assert (cls(res), [cls(a) for a in args]) == hdr[name], name


def test_julia_ccalls_match_the_header():
    """every ccall with a literal symbol in the Julia binding passes the header's argument types."""
    jl_path = os.path.join(os.path.dirname(L.HEADER), "..", "fddgasolver.jl_amd", "julia", "FdgaHIP.jl")
    jl = open(jl_path).read()
    jcls = {"Cint": "i32", "Int32": "i32", "Int64": "i64", "Float64": "f64", "Cdouble": "f64", "Cstring": "str"}
    hdr = _header_signatures()
    calls = re.findall(r"ccall\(\(:(\w+), libfdga\), (\w+), \(([^()]*)\)", jl)
    assert len(calls) >= 20
    for name, ret, args in calls:
        a = [t.strip() for t in args.split(",") if t.strip()]
        got = (jcls[ret], ["ptr" if t.startswith(("Ptr{", "Ref{")) else jcls[t] for t in a])
        assert got == hdr[name], name

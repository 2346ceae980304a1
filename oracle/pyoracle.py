"""ctypes binding of the CPU oracle (oracle/fdga_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product package never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libfdga_oracle.so")

PCH, TCH, ACH = 0, 1, 2
PSP, XSP, DSP = 0, 1, 2
SCPA, FDPA, SCPA_NEW, FDPA_NEW = 0, 1, 2, 3
INF = -(2 ** 31)

GBARE, G0, SIGMA0, G, SIGMA = 0, 1, 2, 3, 4
PI0PP, PI0PH, PIPP, PIPH = 5, 6, 7, 8
LPP, LPH = 9, 10
TAB_FPP, TAB_FPX, TAB_FTP, TAB_FTX = 11, 12, 13, 14
CACHE0 = 32
CACHE_NAMES = ["Gpx", "F0p", "F0a", "F0t", "Gpp", "Ga", "Gt", "Fp", "Fa", "Ft"]
VTX_F, VTX_FBUFF, VTX_FL, VTX_F0 = 0, 1, 2, 3


def vertex_id(vtx, ch, cls):
    return 64 + 16 * vtx + 3 * ch + (cls - 1)


class Params(ctypes.Structure):
    _fields_ = [
        ("nonlocal_", ctypes.c_int32), ("nG", ctypes.c_int32),
        ("nK1", ctypes.c_int32), ("nK2b", ctypes.c_int32), ("nK2f", ctypes.c_int32),
        ("nK3b", ctypes.c_int32), ("nK3f", ctypes.c_int32), ("mPinu", ctypes.c_int32),
        ("LG", ctypes.c_int32), ("L", ctypes.c_int32),
        ("T", ctypes.c_double), ("U", ctypes.c_double),
        ("ref_kind", ctypes.c_int32),
        ("r_nK1", ctypes.c_int32), ("r_nK2b", ctypes.c_int32), ("r_nK2f", ctypes.c_int32),
        ("r_nK3b", ctypes.c_int32), ("r_nK3f", ctypes.c_int32),
        ("tab_nb", ctypes.c_int32), ("tab_nf", ctypes.c_int32),
        ("fdpa_h0", ctypes.c_int32),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i32, i64, dp = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)
        sig = {
            "orc_create": (vp, [ctypes.POINTER(Params)]),
            "orc_destroy": (None, [vp]),
            "orc_size": (i64, [vp, i32]),
            "orc_set": (ctypes.c_int, [vp, i32, dp]),
            "orc_get": (ctypes.c_int, [vp, i32, dp]),
            "orc_set_U": (None, [vp, ctypes.c_double, ctypes.c_double]),
            "orc_siam_bare_green": (None, [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
            "orc_hubbard_bare_green": (None, [vp] + [ctypes.c_double] * 4),
            "orc_init": (None, [vp]),
            "orc_init_sym_grp": (ctypes.c_int, [vp]),
            "orc_sym_size": (ctypes.c_int64, [vp, i32]),
            "orc_sym_classes": (ctypes.c_int, [vp, i32, ctypes.c_void_p, ctypes.c_void_p]),
            "orc_dyson": (None, [vp]),
            "orc_bubbles": (None, [vp]),
            "orc_build_k3_cache": (None, [vp]),
            "orc_bse_l_k2": (None, [vp, i32]),
            "orc_bse_l_k3": (None, [vp, i32]),
            "orc_bse_k1": (None, [vp, i32, i32]),
            "orc_bse_k2": (None, [vp, i32, i32]),
            "orc_bse_k3": (None, [vp, i32, i32]),
            "orc_bse_k1_new": (None, [vp, i32, i32]),
            "orc_bse_k2_new": (None, [vp, i32, i32]),
            "orc_update_F": (None, [vp]),
            "orc_sde": (None, [vp, i32, i32]),
            "orc_iterate": (None, [vp, i32, i32, i32]),
            "orc_build_k3_cache_mfrg": (None, [vp, i32]),
            "orc_mfrg_apply": (None, [vp, dp, dp, i32, i32]),
            "orc_flat_size": (i64, [vp]),
            "orc_flat_size_F": (i64, [vp]),
            "orc_flatten": (None, [vp, dp]),
            "orc_unflatten": (None, [vp, dp]),
            "orc_fixed_point": (None, [vp, dp, dp, i32, i32, i32]),
            "orc_solve": (ctypes.c_int, [vp, i32, i32, ctypes.c_double, ctypes.c_double, i32, i32, dp]),
            "orc_set_threads": (None, [i32]),
            "orc_max_threads": (i32, []),
            "orc_vertex_eval": (None, [vp, i32, i32, i32, i32, i32, i32, i32, i32, dp]),
            "orc_occupation": (ctypes.c_double, [vp, i32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """One solver instance (ParquetSolver or NL_ParquetSolver) of the CPU restatement."""

    def __init__(self, *, nonlocal_=False, nG, nK1, nK2, nK3, mPinu, T, U, LG=1, L=1,
                 ref_kind=0, ref_boxes=None, tab=(1, 1), fdpa_h0=True):
        self.p = Params()
        self.p.nonlocal_ = int(nonlocal_)
        self.p.nG, self.p.nK1 = nG, nK1
        self.p.nK2b, self.p.nK2f = nK2
        self.p.nK3b, self.p.nK3f = nK3
        self.p.mPinu = mPinu
        self.p.LG, self.p.L = LG, L
        self.p.T, self.p.U = T, U
        self.p.ref_kind = ref_kind
        rb = ref_boxes or (1, (1, 1), (1, 1))
        self.p.r_nK1 = rb[0]
        self.p.r_nK2b, self.p.r_nK2f = rb[1]
        self.p.r_nK3b, self.p.r_nK3f = rb[2]
        self.p.tab_nb, self.p.tab_nf = tab
        self.p.fdpa_h0 = int(fdpa_h0)
        self.lib = lib()
        self.h = self.lib.orc_create(ctypes.byref(self.p))
        self.T = T

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def get(self, which):
        n = self.lib.orc_size(self.h, which)
        if n < 0:
            raise KeyError(which)
        out = np.empty(n, dtype=np.complex128)
        self.lib.orc_get(self.h, which, _dp(out.view(np.float64)))
        return out

    def set(self, which, data):
        n = self.lib.orc_size(self.h, which)
        a = np.ascontiguousarray(np.asarray(data, dtype=np.complex128).reshape(-1, order="F"))
        if a.size != n:
            raise ValueError(f"array {which}: expected {n} elements, got {a.size}")
        self.lib.orc_set(self.h, which, _dp(a.view(np.float64)))

    def vertex(self, vtx, ch, cls):
        return self.get(vertex_id(vtx, ch, cls))

    def set_vertex(self, vtx, ch, cls, data):
        self.set(vertex_id(vtx, ch, cls), data)

    # model inputs
    def siam(self, e, Delta, D=None):
        self.lib.orc_siam_bare_green(self.h, e, Delta, -1.0 if D is None else D)

    def hubbard(self, mu, t1, t2=0.0, t3=0.0):
        self.lib.orc_hubbard_bare_green(self.h, mu, t1, t2, t3)

    def init(self):
        self.lib.orc_init(self.h)

    def init_sym_grp(self):
        assert self.lib.orc_init_sym_grp(self.h) == 0

    def sym_classes(self, gid):
        """(representative, operation) per element of an installed group (ids: fdga_oracle.h)."""
        n = self.lib.orc_sym_size(self.h, gid)
        if n < 0:
            raise KeyError(gid)
        rep = np.empty(n, dtype=np.int64)
        op = np.empty(n, dtype=np.int32)
        assert self.lib.orc_sym_classes(self.h, gid, rep.ctypes.data, op.ctypes.data) == 0
        return rep, op

    # stages
    def dyson(self): self.lib.orc_dyson(self.h)
    def bubbles(self): self.lib.orc_bubbles(self.h)
    def build_k3_cache(self): self.lib.orc_build_k3_cache(self.h)
    def bse_l_k2(self, ch): self.lib.orc_bse_l_k2(self.h, ch)
    def bse_l_k3(self, ch): self.lib.orc_bse_l_k3(self.h, ch)
    def bse_k1(self, ch, is_mfrg=False): self.lib.orc_bse_k1(self.h, ch, int(is_mfrg))
    def bse_k2(self, ch, is_mfrg=False): self.lib.orc_bse_k2(self.h, ch, int(is_mfrg))
    def bse_k3(self, ch, is_mfrg=False): self.lib.orc_bse_k3(self.h, ch, int(is_mfrg))
    def bse_k1_new(self, ch, is_mfrg=False): self.lib.orc_bse_k1_new(self.h, ch, int(is_mfrg))
    def bse_k2_new(self, ch, is_mfrg=False): self.lib.orc_bse_k2_new(self.h, ch, int(is_mfrg))
    def build_k3_cache_mfrg(self, is_first=True): self.lib.orc_build_k3_cache_mfrg(self.h, int(is_first))
    def update_F(self): self.lib.orc_update_F(self.h)
    def sde(self, strategy, include_hartree=True): self.lib.orc_sde(self.h, strategy, int(include_hartree))

    def iterate(self, strategy, update_sigma=True, compute_hartree=True):
        self.lib.orc_iterate(self.h, strategy, int(update_sigma), int(compute_hartree))

    def mfrg_apply(self, x, strategy=FDPA, is_first=True):
        """mfRGLinearMap(S, strategy) * x on the vertex-only vector x."""
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.empty_like(x)
        self.lib.orc_mfrg_apply(self.h, _dp(y.view(np.float64)), _dp(x.view(np.float64)), strategy, int(is_first))
        return y

    def flat_size_F(self):
        return self.lib.orc_flat_size_F(self.h)

    def flatten(self):
        n = self.lib.orc_flat_size(self.h)
        x = np.empty(n, dtype=np.complex128)
        self.lib.orc_flatten(self.h, _dp(x.view(np.float64)))
        return x

    def unflatten(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        self.lib.orc_unflatten(self.h, _dp(x.view(np.float64)))

    def fixed_point(self, x, strategy, update_sigma=True, compute_hartree=True):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        R = np.empty_like(x)
        self.lib.orc_fixed_point(self.h, _dp(R.view(np.float64)), _dp(x.view(np.float64)), strategy,
                                 int(update_sigma), int(compute_hartree))
        return R

    def solve(self, strategy, maxiter=100, tol=1e-4, beta=0.85, mem=8, update_sigma=True):
        res = ctypes.c_double(0.0)
        it = self.lib.orc_solve(self.h, strategy, maxiter, tol, beta, mem, int(update_sigma), ctypes.byref(res))
        return it, res.value

    def eval(self, vtx, m, n, np_, iP, ch, sp, flags=15):
        out = (ctypes.c_double * 2)()
        self.lib.orc_vertex_eval(self.h, vtx, m, n, np_, iP, ch, sp, flags, out)
        return complex(out[0], out[1])

    def occupation(self, which=G):
        return self.lib.orc_occupation(self.h, which)


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops (results do not depend on the count)."""
    lib().orc_set_threads(int(n))


def max_threads():
    return lib().orc_max_threads()

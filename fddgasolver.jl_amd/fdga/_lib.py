"""ctypes binding of libfdga.so (include/fdga.h).

The library is the only compute path: if it is missing or no HIP device is present the
calls fail with an exception; nothing falls back to a CPU implementation.
"""
import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libfdga.so")
CSRC = os.path.join(PKG_ROOT, "csrc")
HEADER = os.path.join(os.path.dirname(PKG_ROOT), "include", "fdga.h")

# array identifiers (include/fdga.h)
GBARE, G0, SIGMA0, G, SIGMA = 0, 1, 2, 3, 4
PI0PP, PI0PH, PIPP, PIPH = 5, 6, 7, 8
LPP, LPH = 9, 10
TAB_FPP, TAB_FPX, TAB_FTP, TAB_FTX = 11, 12, 13, 14
CACHE0 = 32
PCH, TCH, ACH = 0, 1, 2
SCPA, FDPA, SCPA_NEW, FDPA_NEW = 0, 1, 2, 3
FP64, FP32 = 0, 1
TILE_64x64, TILE_32x32, TILE_64x32, TILE_32x64 = 0, 1, 2, 3


def vertex_id(vtx, ch, cls):
    """vtx: 0 F, 1 Fbuff, 2 FL, 3 F0 (local reference Vertex)."""
    return 64 + 16 * vtx + 3 * ch + (cls - 1)


class Dims(ctypes.Structure):
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
        ("fdpa_h0", ctypes.c_int32), ("precision", ctypes.c_int32),
    ]


_vp, _i32, _i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
_dp = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "fdga_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(Dims)]),
    "fdga_destroy": (ctypes.c_int, [_vp]),
    "fdga_last_error": (ctypes.c_char_p, [_vp]),
    "fdga_size": (_i64, [_vp, _i32]),
    "fdga_set": (ctypes.c_int, [_vp, _i32, _dp]),
    "fdga_get": (ctypes.c_int, [_vp, _i32, _dp]),
    "fdga_init": (ctypes.c_int, [_vp]),
    "fdga_init_sym_grp": (ctypes.c_int, [_vp]),
    "fdga_gemm_blocks": (_i64, [_vp, _i32]),
    "fdga_gemm_flops": (ctypes.c_double, [_vp, _i32]),
    "fdga_gemm_tile": (ctypes.c_int, [_vp, _i32]),
    "fdga_set_gemm_tile": (ctypes.c_int, [_vp, _i32, _i32]),
    "fdga_gemm_grid": (ctypes.c_int, [_vp, _i32]),
    "fdga_set_gemm_grid": (ctypes.c_int, [_vp, _i32, _i32]),
    "fdga_autotune": (ctypes.c_int, [_vp, _i32]),
    "fdga_set_graph": (ctypes.c_int, [_vp, _i32]),
    "fdga_reset_sym_grp": (ctypes.c_int, [_vp]),
    "fdga_dyson": (ctypes.c_int, [_vp]),
    "fdga_bubbles": (ctypes.c_int, [_vp]),
    "fdga_build_k3_cache": (ctypes.c_int, [_vp]),
    "fdga_bse_l_k2": (ctypes.c_int, [_vp, _i32]),
    "fdga_bse_l_k3": (ctypes.c_int, [_vp, _i32]),
    "fdga_bse_k1": (ctypes.c_int, [_vp, _i32, _i32]),
    "fdga_bse_k2": (ctypes.c_int, [_vp, _i32, _i32]),
    "fdga_bse_k3": (ctypes.c_int, [_vp, _i32, _i32]),
    "fdga_bse_k1_new": (ctypes.c_int, [_vp, _i32, _i32]),
    "fdga_bse_k2_new": (ctypes.c_int, [_vp, _i32, _i32]),
    "fdga_build_k3_cache_mfrg": (ctypes.c_int, [_vp, _i32]),
    "fdga_update_F": (ctypes.c_int, [_vp]),
    "fdga_sde": (ctypes.c_int, [_vp, _i32, _i32]),
    "fdga_iterate": (ctypes.c_int, [_vp, _i32, _i32, _i32]),
    "fdga_flat_size": (_i64, [_vp]),
    "fdga_flat_size_F": (_i64, [_vp]),
    "fdga_flatten": (ctypes.c_int, [_vp, _dp]),
    "fdga_unflatten": (ctypes.c_int, [_vp, _dp]),
    "fdga_fixed_point": (ctypes.c_int, [_vp, _dp, _dp, _i32, _i32, _i32]),
    "fdga_fixed_point_noalloc": (ctypes.c_int, [_vp, _dp, _dp, _i32, _i32, _i32]),
    "fdga_mfrg_apply": (ctypes.c_int, [_vp, _dp, _dp, _i32, _i32]),
    "fdga_bse_k3_jvp": (ctypes.c_int, [_vp, _vp]),
    "fdga_bse_k3_vjp": (ctypes.c_int, [_vp, _vp]),
    "fdga_k3_gather_map": (_i64, [ctypes.POINTER(Dims), ctypes.POINTER(_i64), ctypes.POINTER(_i32),
                                  ctypes.POINTER(ctypes.c_float)]),
    "fdga_sync": (ctypes.c_int, [_vp]),
    "fdga_time_iterations": (ctypes.c_int, [_vp, _i32, _i32, _dp]),
    "fdga_stage_ms": (ctypes.c_int, [_vp, _i32, _dp]),
    "fdga_set_stage_timing": (ctypes.c_int, [_vp, _i32]),
    "fdga_selftest_zgemm": (ctypes.c_int, [_i32, _i32, _i32, _i32, _dp, _dp, _dp]),
    "fdga_selftest_zgemm_tiles": (ctypes.c_int, [_i32, _i32, _i32, _i32, _i32, _i64, _dp, _dp, _dp]),
    "fdga_selftest_zgemm_batched": (ctypes.c_int, [_i32, _i32, _i32, _i32, _i32, _i32, _i64, _dp, _dp, _dp]),
    "fdga_group_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p), _i32]),
    "fdga_group_destroy": (ctypes.c_int, [_vp]),
    "fdga_group_last_error": (ctypes.c_char_p, [_vp]),
    "fdga_group_iterate": (ctypes.c_int, [_vp, _i32, _i32]),
    "fdga_group_sync": (ctypes.c_int, [_vp]),
    "fdga_group_time_iterations": (ctypes.c_int, [_vp, _i32, _i32, _dp]),
    "fdga_group_set_graph": (ctypes.c_int, [_vp, _i32]),
    "fdga_group_set_timing": (ctypes.c_int, [_vp, _i32]),
    "fdga_group_gemm_ms": (ctypes.c_int, [_vp, _i32, _dp]),
    "fdga_group_tiles": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "fdga_group_set_tiles": (ctypes.c_int, [_vp, _i32, _i32, _i64]),
    "fdga_group_autotune": (ctypes.c_int, [_vp, _i32]),
}

_lib = None


def build(verbose=False):
    """Compile libfdga.so for gfx950 in-tree (hipcc cross-compiles without a GPU)."""
    out = subprocess.run(["make", "-B", "-C", CSRC], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("building libfdga.so failed:\n" + out.stdout + out.stderr)
    if verbose:
        print(out.stdout)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib

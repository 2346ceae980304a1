"""The batched complex GEMM engine against numpy: FP64 MFMA to 1e-13, FP32 MFMA (operands rounded
to FP32, FP32 accumulation) to 5e-6 relative to the largest entry."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@pytest.mark.parametrize("precision,tol", [(0, 1e-13), (1, 5e-6)])
@pytest.mark.parametrize("M,N,K", [(16, 16, 4), (6, 6, 6), (64, 64, 16), (37, 70, 29), (128, 128, 128)])
def test_zgemm_engine(M, N, K, precision, tol):
    import fdga
    lib = fdga.load()
    rng = np.random.default_rng(M * 1000 + N + K)
    A = np.asfortranarray(rng.standard_normal((M, K)) + 1j * rng.standard_normal((M, K)))
    B = np.asfortranarray(rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N)))
    a = np.ascontiguousarray(A.ravel(order="F"))
    b = np.ascontiguousarray(B.ravel(order="F"))
    c = np.zeros(M * N, dtype=np.complex128)
    rc = lib.fdga_selftest_zgemm(M, N, K, precision, _dp(a.view(np.float64)), _dp(b.view(np.float64)),
                                 _dp(c.view(np.float64)))
    assert rc == 0
    Cg = c.reshape(M, N, order="F")
    ref = A @ B
    err = np.max(np.abs(Cg - ref)) / np.max(np.abs(ref))
    if err > tol and M % 16 == 0 and N % 16 == 0:
        # diagnose the accumulator lane layout: compare against 16x16 block transposes
        refT = ref.reshape(M // 16, 16, N // 16, 16).transpose(0, 3, 2, 1).reshape(M, N)
        print("block-transposed error", np.max(np.abs(Cg - refT)) / np.max(np.abs(ref)))
        print(np.round(Cg[:16, :16].real - ref[:16, :16].real, 3))
    assert err < tol, err


@pytest.mark.parametrize("precision,tol", [(0, 1e-13), (1, 5e-6)])
@pytest.mark.parametrize("M,N,K", [(72, 88, 56), (96, 96, 96), (20, 130, 40)])
def test_zgemm_tile_shapes_and_persistent_grids(M, N, K, precision, tol):
    """Every tile shape (workgroups of 2x2, 1x1, 2x1, 1x2 waves), one workgroup per tile or a
    persistent grid of 8 workgroups looping over the tiles: the same bits as the 64x64 launch
    (each element accumulates over k in the same order) and numpy's product within `tol`."""
    import fdga
    lib = fdga.load()
    rng = np.random.default_rng(M + 7 * N + 13 * K)
    A = np.asfortranarray(rng.standard_normal((M, K)) + 1j * rng.standard_normal((M, K)))
    B = np.asfortranarray(rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N)))
    a = np.ascontiguousarray(A.ravel(order="F"))
    b = np.ascontiguousarray(B.ravel(order="F"))
    base = np.zeros(M * N, dtype=np.complex128)
    assert lib.fdga_selftest_zgemm(M, N, K, precision, _dp(a.view(np.float64)), _dp(b.view(np.float64)),
                                   _dp(base.view(np.float64))) == 0
    ref = A @ B
    for tile in range(4):
        for grid in (0, 8):
            c = np.zeros(M * N, dtype=np.complex128)
            rc = lib.fdga_selftest_zgemm_tiles(M, N, K, precision, tile, grid, _dp(a.view(np.float64)),
                                               _dp(b.view(np.float64)), _dp(c.view(np.float64)))
            assert rc == 0
            np.testing.assert_array_equal(c, base, err_msg=f"tile {tile} grid {grid}")
            err = np.max(np.abs(c.reshape(M, N, order="F") - ref)) / np.max(np.abs(ref))
            assert err < tol, (tile, grid, err)


@pytest.mark.parametrize("precision,tol", [(0, 1e-13), (1, 5e-6)])
def test_zgemm_batched(precision, tol):
    """A batch of independent products (the shape of the BSE contractions: many small products),
    fewer and more than 8 batches, every tile shape."""
    import fdga
    lib = fdga.load()
    M = N = K = 40
    for nbatch in (3, 19):
        rng = np.random.default_rng(nbatch)
        A = rng.standard_normal((nbatch, K, M)) + 1j * rng.standard_normal((nbatch, K, M))   # [b, k, i]: column-major A[b]
        B = rng.standard_normal((nbatch, N, K)) + 1j * rng.standard_normal((nbatch, N, K))
        ref = np.einsum("bki,bjk->bji", A, B)   # C[b] column-major: [b, j, i]
        for tile in range(4):
            c = np.zeros(nbatch * M * N, dtype=np.complex128)
            a = np.ascontiguousarray(A.ravel())
            b = np.ascontiguousarray(B.ravel())
            rc = lib.fdga_selftest_zgemm_batched(M, N, K, nbatch, precision, tile, 0, _dp(a.view(np.float64)),
                                                 _dp(b.view(np.float64)), _dp(c.view(np.float64)))
            assert rc == 0
            err = np.max(np.abs(c.reshape(nbatch, N, M) - ref)) / np.max(np.abs(ref))
            assert err < tol, (nbatch, tile, err)

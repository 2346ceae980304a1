"""Tile shapes and persistent grids of the single-solver contraction launches and the autotuner: every
shape and grid gives the same bits (each element accumulates over k in the same order), with the
symmetry block lists of that shape on the local solver; the autotuner leaves the solver state as it
was."""
import numpy as np
import pytest

from helpers import make_pair, relerr

pytestmark = pytest.mark.gpu
TILES = ("64x64", "32x32", "64x32", "32x64")


@pytest.mark.parametrize("case,sym", [("local_vertexref", True), ("local_ref", False), ("nl_L2", False)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_every_tile_shape_same_iteration(case, sym, precision):
    import fdga
    ref, _ = make_pair(case, seed=6, precision=precision)
    if sym:
        ref.init_sym_grp()
    x0 = fdga.flatten(ref)
    fdga.iterate_solver(ref, "fdPA")
    want = fdga.flatten(ref)
    S, _ = make_pair(case, seed=6, precision=precision)
    if sym:
        S.init_sym_grp()
    for tl in TILES:
        for tk in TILES:
            fdga.unflatten(S, x0)
            S.set_gemm_tile(True, tl)
            S.set_gemm_tile(False, tk)
            assert S.gemm_tiles() == (tl, tk)
            fdga.iterate_solver(S, "fdPA")
            got = fdga.flatten(S)
            assert relerr(got, want) == 0.0, (tl, tk, relerr(got, want))


@pytest.mark.parametrize("case", ["nl_L16", "local_vertexref"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_persistent_gemm_grid_same_iteration(case, precision):
    """A persistent grid for the contraction launches (workgroups looping over the output tiles, the
    hipGraph replay included) gives the same bits as one workgroup per tile; invalid counts are refused."""
    import fdga
    ref, _ = make_pair(case, seed=7, precision=precision)
    x0 = fdga.flatten(ref)
    for _ in range(2):
        fdga.iterate_solver(ref, "fdPA")
    want = fdga.flatten(ref)
    S, _ = make_pair(case, seed=7, precision=precision)
    with pytest.raises(fdga.FdgaError, match="multiple of 8"):
        S.set_gemm_grid(True, 100)
    for graph in (False, True):
        S.set_graph(graph)
        for g in (8, 192, 256, 512, 0):
            fdga.unflatten(S, x0)
            S.set_gemm_grid(True, g)
            S.set_gemm_grid(False, g)
            assert S.gemm_grids() == (g, g)
            for _ in range(2):
                fdga.iterate_solver(S, "fdPA")
            S.sync()
            got = fdga.flatten(S)
            assert relerr(got, want) == 0.0, (graph, g, relerr(got, want))
    ref.close()
    S.close()


@pytest.mark.parametrize("strategy", ["scPA", "fdPA"])
def test_autotune_leaves_state(strategy):
    import fdga
    A, _ = make_pair("local_vertexref", seed=8)
    B, _ = make_pair("local_vertexref", seed=8)
    for S in (A, B):
        S.init_sym_grp()
        fdga.iterate_solver(S, strategy)
    B.autotune(reps=1)
    assert all(t in TILES for t in B.gemm_tiles())
    assert B.gemm_flops(True) > 0 and B.gemm_flops(False) > 0
    for it in range(2):
        fdga.iterate_solver(A, strategy)
        fdga.iterate_solver(B, strategy)
        assert relerr(fdga.flatten(B), fdga.flatten(A)) == 0.0, it

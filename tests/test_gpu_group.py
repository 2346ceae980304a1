"""Groups of solvers iterated together (parameter sweeps): each member's stages on its own stream,
the BSE_L_K3!/BSE_K3! contractions of all members in one grouped GEMM launch each.  Every member
must end exactly where iterate_solver on the member alone ends (same kernels, and the grouped
GEMM accumulates every element in the same order whatever the tile shape or grid), under every
tile shape and persistent grid, with the hipGraph replay and after the autotuner; and every grouped
member equals the CPU oracle run on the same inputs (1e-12).  The full-size sweep is in
tests/test_gpu_sweep.py."""
import numpy as np
import pytest

from helpers import BOX, make_pair, relerr

pytestmark = pytest.mark.gpu

# lattice members with uneven boxes and lattices (K3 boxes (3,3), (3,2), (2,4); L = 2, 3, 16)
NL_MEMBERS = [("nl_L2", BOX), ("nl_L3", dict(BOX, nK3=(3, 2))), ("nl_L16", dict(BOX, nK3=(2, 4)))]
LOCAL_MEMBERS = [("local_ref", BOX), ("local_vertexref", dict(BOX, nK3=(2, 4)))]


def _pairs(members, seed, precision="fp64"):
    solo = [make_pair(c, seed=seed + k, box=b, precision=precision)[0] for k, (c, b) in enumerate(members)]
    grp = [make_pair(c, seed=seed + k, box=b, precision=precision)[0] for k, (c, b) in enumerate(members)]
    return solo, grp


def _assert_same(solo, grp, what):
    import fdga
    for k, (a, b) in enumerate(zip(solo, grp)):
        x, y = fdga.flatten(a), fdga.flatten(b)
        assert relerr(y, x) == 0.0, f"{what} member {k}: {relerr(y, x):.3e}"


@pytest.mark.parametrize("strategy", ["fdPA", "scPA"])
@pytest.mark.parametrize("members", ["nl", "local"])
def test_group_iteration_equals_members_alone(members, strategy):
    import fdga
    import pyoracle as po
    mem = NL_MEMBERS if members == "nl" else LOCAL_MEMBERS
    solo, grp = _pairs(mem, seed=11)
    orcs = [make_pair(c, seed=11 + k, box=b)[1] for k, (c, b) in enumerate(mem)]
    G = fdga.SolverGroup(grp)
    try:
        for it in range(2):
            for S in solo:
                fdga.iterate_solver(S, strategy)
            for O in orcs:
                O.iterate(po.SCPA if strategy == "scPA" else po.FDPA)
            G.iterate(strategy)
            G.sync()
            _assert_same(solo, grp, f"{members} {strategy} iteration {it}")
            for k, (S, O) in enumerate(zip(grp, orcs)):
                e = relerr(fdga.flatten(S), O.flatten())
                assert e < 1e-12, f"{members} {strategy} iteration {it} member {k} vs oracle: {e:.3e}"
    finally:
        G.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_group_every_tile_shape_and_grid(precision):
    import fdga
    solo, grp = _pairs(NL_MEMBERS, seed=21, precision=precision)
    G = fdga.SolverGroup(grp)
    try:
        it = 0
        for tile in ("64x64", "32x32", "64x32", "32x64"):
            for grid in (0, 8):
                G.set_tiles(0, tile, grid)
                G.set_tiles(1, tile, grid)
                assert G.tiles(1)[:2] == (tile, grid)
                for S in solo:
                    fdga.iterate_solver(S, "fdPA")
                G.iterate("fdPA")
                G.sync()
                _assert_same(solo, grp, f"{precision} tile {tile} grid {grid} iteration {it}")
                it += 1
    finally:
        G.close()


def test_group_matches_members_at_64_frequencies():
    """FP32 lattice members whose K3 box has Nν3 = 64, the headline's: the single solver's
    momentum-pipelined BSE_L_K3! kernel gives the bits of the grouped GEMM engine (non-zero FL: a local
    reference vertex with RefVertex tables)."""
    import fdga
    box = dict(nG=48, nK1=32, nK2=(4, 32), nK3=(2, 32), mPinu=1, T=0.3, U=1.3)
    mem = [("nl_L16_ref", box), ("nl_L16_ref", dict(box, nK3=(3, 32)))]
    solo, grp = _pairs(mem, seed=31, precision="fp32")
    solo[1].set_gemm_grid(True, 8)   # the batched kernel's workgroups looping over its tasks
    G = fdga.SolverGroup(grp)
    try:
        for it in range(2):
            for S in solo:
                fdga.iterate_solver(S, "fdPA")
            G.iterate("fdPA")
            G.sync()
            _assert_same(solo, grp, f"Nν3 = 64 iteration {it}")
        for S in solo:
            assert max(np.abs(S.vertex(2, ch, 3)).max() for ch in range(3)) > 1e-6
    finally:
        G.close()


def test_group_graph_replay_and_autotune():
    """Eager first iteration (caches the reference SDE), then hipGraph replays with eager work in
    between; then the autotuner (which must leave the state alone) and more replays."""
    import fdga
    solo, grp = _pairs(NL_MEMBERS, seed=31)
    G = fdga.SolverGroup(grp)
    try:
        G.set_graph(True)
        for it in range(3):
            for S in solo:
                fdga.iterate_solver(S, "fdPA")
            G.iterate("fdPA")
            G.sync()
            _assert_same(solo, grp, f"graph replay {it}")
        G.autotune(reps=1)
        for which in (0, 1):
            tile, grid, computed = G.tiles(which)
            assert tile in G.TILES and computed > 0
        _assert_same(solo, grp, "after autotune")
        for it in range(2):
            for S in solo:
                fdga.iterate_solver(S, "fdPA")
            G.iterate("fdPA")
            G.sync()
            _assert_same(solo, grp, f"replay after autotune {it}")
    finally:
        G.close()


def test_group_errors():
    import fdga
    a = make_pair("nl_L2", seed=1)[0]
    b = make_pair("local_ref", seed=1)[0]
    with pytest.raises(fdga.FdgaError, match="all local or all lattice"):
        fdga.SolverGroup([a, b])
    c = make_pair("nl_L2", seed=1, precision="fp32")[0]
    with pytest.raises(fdga.FdgaError, match="precision"):
        fdga.SolverGroup([a, c])
    with pytest.raises(fdga.FdgaError, match="twice"):
        fdga.SolverGroup([a, a])
    d = make_pair("local_vertexref", seed=2)[0]
    d.init_sym_grp()
    with pytest.raises(fdga.FdgaError, match="symmetry"):
        fdga.SolverGroup([b, d])
    G = fdga.SolverGroup([a])
    with pytest.raises(fdga.FdgaError, match="tile"):
        G._check(G.lib.fdga_group_set_tiles(G.h, 1, 9, 0))
    with pytest.raises(fdga.FdgaError, match="multiple of 8"):
        G.set_tiles(1, "32x32", 12)
    G.close()

"""GPU parity of the other BSE forms and boundary parameters against the CPU oracle (1e-12 relative):
  * BSE_K1!/BSE_K2!/BSE_K3! with is_mfRG = Val(true) (src/BSEa/BSEa_K{1,2,3}.jl; lattice
    src/nonlocal/BSEa/BSEa_K{1,2,3}.jl) and BSE_K1_new!/BSE_K2_new! with both is_mfRG values,
    with and without symmetry groups, on random non-symmetric inputs (FL with random K2/K3);
  * build_K3_cache_mfRG! (src/nonlocal_2/build_K3_cache.jl:97-164) and mfRGLinearMap * x
    (src/mfRG.jl:34-89) for :fdPA and :fdPA_new, first and later products;
  * iterate_solver! with :scPA_new / :fdPA_new (src/solve.jl:26-46), eagerly and replayed from a
    hipGraph, and with compute_Hartree = false (src/solve.jl:7,99);
  * fixed_point! with update_Σ = false on vertex-only vectors (src/solve.jl:131-139).
The oracle side is pinned by tests/test_oracle_pins.py."""
import numpy as np
import pytest

import pyoracle as po
from helpers import CASES, make_pair, rand_c, relerr

pytestmark = pytest.mark.gpu

import fdga  # noqa: E402

TOL = 1e-12
LATTICE = [c for c in CASES if CASES[c][0]]


def rel(a, b):
    """relative to max|b|; absolute when b vanishes (the local is_mfRG BSE_K1! has a zero t channel)"""
    s = np.max(np.abs(b)) if np.size(b) else 0.0
    d = float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) if np.size(b) else 0.0
    return d / s if s > 0 else d


def random_FL(S, O, seed):
    rng = np.random.default_rng(seed)
    for ch in range(3):
        for cls in (2, 3):
            v = rand_c(rng, O.vertex(po.VTX_FL, ch, cls).size, 0.2)
            S.set_vertex(2, ch, cls, v)
            O.set_vertex(po.VTX_FL, ch, cls, v)


@pytest.mark.parametrize("sym", [False, True], ids=["dense", "groups"])
@pytest.mark.parametrize("case", list(CASES))
def test_k1_k2_forms(case, sym):
    print(f"forms: {case} sym={sym}", flush=True)
    S, O = make_pair(case, seed=31)
    random_FL(S, O, 32)
    if sym:
        S.init_sym_grp()
        O.init_sym_grp()
    stages = [
        ("K1 mfRG", 1, lambda ch: fdga.BSE_K1(S, ch, True), lambda ch: O.bse_k1(ch, True)),
        ("K2 mfRG", 2, lambda ch: fdga.BSE_K2(S, ch, True), lambda ch: O.bse_k2(ch, True)),
        ("K1_new", 1, lambda ch: fdga.BSE_K1_new(S, ch), lambda ch: O.bse_k1_new(ch)),
        ("K2_new", 2, lambda ch: fdga.BSE_K2_new(S, ch), lambda ch: O.bse_k2_new(ch)),
        ("K1_new mfRG", 1, lambda ch: fdga.BSE_K1_new(S, ch, True), lambda ch: O.bse_k1_new(ch, True)),
        ("K2_new mfRG", 2, lambda ch: fdga.BSE_K2_new(S, ch, True), lambda ch: O.bse_k2_new(ch, True)),
    ]
    for name, cls, dev, orc in stages:
        for ch in (po.PCH, po.ACH, po.TCH):
            dev(ch)
            orc(ch)
        for ch in range(3):
            e = rel(S.vertex(1, ch, cls), O.vertex(po.VTX_FBUFF, ch, cls))
            assert e < TOL, (name, ch, e)
    S.close()


@pytest.mark.parametrize("sym", [False, True], ids=["dense", "groups"])
@pytest.mark.parametrize("case", list(CASES))
def test_k3_mfrg(case, sym):
    print(f"K3 mfRG: {case} sym={sym}", flush=True)
    nl = CASES[case][0]
    if nl and sym:
        pytest.skip("build_K3_cache_mfRG! is refused under lattice symmetry groups (momentum-dependent class fill)")
    S, O = make_pair(case, seed=33)
    random_FL(S, O, 34)
    if sym:
        S.init_sym_grp()
        O.init_sym_grp()
    fdga.build_K3_cache(S)
    O.build_k3_cache()
    if nl:
        with pytest.raises(fdga.FdgaError, match="build_K3_cache_mfRG"):
            fdga.BSE_K3(S, po.PCH, True)
        fdga.build_K3_cache_mfRG(S, True)
        O.build_k3_cache_mfrg(True)
        for c in range(10):
            e = relerr(S.get(fdga._lib.CACHE0 + c), O.get(po.CACHE0 + c))
            assert e < TOL, (po.CACHE_NAMES[c], e)
    for ch in (po.PCH, po.ACH, po.TCH):
        fdga.BSE_K3(S, ch, True)
        O.bse_k3(ch, True)
    for ch in range(3):
        e = rel(S.vertex(1, ch, 3), O.vertex(po.VTX_FBUFF, ch, 3))
        assert e < TOL, (ch, e)
    S.close()


def test_mfrg_cache_refused_for_local_solver():
    S, _ = make_pair("local_ref", seed=1)
    with pytest.raises(fdga.FdgaError, match="no method for ParquetSolver"):
        fdga.build_K3_cache_mfRG(S, True)
    with pytest.raises(fdga.FdgaError, match="no method for ParquetSolver"):
        fdga.mfRGLinearMap(S, "fdPA").matvec(np.zeros(S.lib.fdga_flat_size_F(S.h), complex))
    S.close()


@pytest.mark.parametrize("strategy", ["fdPA", "fdPA_new"])
@pytest.mark.parametrize("case", LATTICE)
def test_mfrg_linear_map(case, strategy):
    print(f"mfRGLinearMap: {case} {strategy}", flush=True)
    S, O = make_pair(case, seed=35)
    random_FL(S, O, 36)
    # the map keeps the F0 caches of the last build_K3_cache! (iterate_solver! precedes it in the reference)
    fdga.build_K3_cache(S)
    O.build_k3_cache()
    rng = np.random.default_rng(37)
    n = O.flat_size_F()
    A = fdga.mfRGLinearMap(S, strategy)
    ost = po.FDPA if strategy == "fdPA" else po.FDPA_NEW
    for k in range(2):   # the first product builds the F caches (is_first_iteration), the second reuses them
        x = rand_c(rng, n, 0.3)
        y = A @ x
        yo = O.mfrg_apply(x, ost, k == 0)
        e = relerr(y, yo)
        assert e < 1e-11, (k, e)
    S.close()


@pytest.mark.parametrize("strategy", ["scPA_new", "fdPA_new"])
@pytest.mark.parametrize("case", list(CASES))
def test_iterate_new_strategies(case, strategy):
    print(f"iterate {strategy}: {case}", flush=True)
    S, O = make_pair(case, seed=38)
    ost = po.SCPA_NEW if strategy == "scPA_new" else po.FDPA_NEW
    for it in range(2):
        fdga.iterate_solver(S, strategy)
        O.iterate(ost)
        e = relerr(fdga.flatten(S), O.flatten())
        assert e < TOL, (it, e)
    S.close()


def test_new_strategy_graph_replay():
    """:fdPA_new replayed from the captured hipGraph of one iteration equals the eager oracle iterations."""
    S, O = make_pair("local_vertexref", seed=39)
    S.set_graph(True)
    for it in range(4):
        fdga.iterate_solver(S, "fdPA_new")
        O.iterate(po.FDPA_NEW)
        S.sync()
        e = relerr(fdga.flatten(S), O.flatten())
        assert e < TOL, (it, e)
    S.close()


@pytest.mark.parametrize("strategy", ["scPA", "fdPA"])
@pytest.mark.parametrize("case", ["local_vertexref", "nl_L2"])
def test_compute_hartree_false(case, strategy):
    S, O = make_pair(case, seed=40)
    st = po.SCPA if strategy == "scPA" else po.FDPA
    for it in range(2):
        fdga.iterate_solver(S, strategy, compute_Hartree=False)
        O.iterate(st, compute_hartree=False)
        assert relerr(fdga.flatten(S), O.flatten()) < TOL, it
    S.close()


@pytest.mark.parametrize("case", ["local_ref", "nl_L3"])
def test_fixed_point_update_sigma_false(case):
    S, O = make_pair(case, seed=41)
    n = O.flat_size_F()
    x = O.flatten()[:n] * 0.9
    R = np.empty(n, dtype=np.complex128)
    fdga.fixed_point(R, x, S, "fdPA", update_Σ=False)
    Ro = O.fixed_point(x, po.FDPA, update_sigma=False)
    assert relerr(R, Ro) < TOL
    # Σ, G and Π are left as they were: a second call from the same x gives the same R
    R2 = np.empty_like(R)
    fdga.fixed_point(R2, x, S, "fdPA", update_Σ=False)
    assert relerr(R2, R) < TOL
    S.close()

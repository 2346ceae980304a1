"""Parity of the lattice (s-wave) path at the full size of the lattice workload through an exact
reduction the domain fixes: with k-independent Green functions (G(k) = G_loc for every k, on the
16 x 16 lattice) G(R) vanishes away from R = 0, so Π(Ω, ν, q) does not depend on q, neither do the
vertex and Σ, and every momentum component of NL_ParquetSolver must equal ParquetSolver with the same
boxes, the RefVertex(U) reference and the same (P-independent) starting vertex.  (With a local Vertex
as reference, or a non-zero G0, the two solvers' BSE_L_K2! are different formulas in the reference --
the lattice one sums over the K2 mesh only, src/BSEa/BSEa_K2.jl:1-40 vs src/nonlocal/BSEa/BSEa_K2.jl:1-41
-- so the reduction is taken with RefVertex(U) and G0 = Σ0 = 0, as in the oracle's own reduction test
at L = 1, 2, 3.)  The lattice path (fused real-space bubble with the L = 16 FFT, band-stored
weights, K2 / L_K2 tables and GEMMs, the lattice SDE, the P-averages, the lattice BSE_K3! loaders)
shares no kernel with the local path's windowed gathers, which the oracle pins at small sizes.
Boxes of the lattice workload: nK1 = 64, nK2 = (32, 32), nK3 = (16, 32), mΠν = 32, nG = 256,
T = 0.2, U = 2.  FP64 to 1e-10, the FP32 contraction to 1e-5 (the two paths round different but
equal operand sums to FP32)."""
import numpy as np
import pytest

from helpers import relerr

pytestmark = pytest.mark.gpu

W = dict(L=16, T=0.2, U=2.0, nK1=64, nK2=(32, 32), nK3=(16, 32), mPinu=32, nG=256)
# the largest point of the batched sweep: 24 x 24 lattice (the 3 x 8 FFT of the fused bubble) and
# Nν3 = 96 > 64 (the BSE_K3! right-operand builder with two ω per lane)
W24 = dict(L=24, T=0.1, U=2.9, nK1=96, nK2=(48, 48), nK3=(24, 48), mPinu=1, nG=192)


def _pair(precision, w=W, seed=3):
    import fdga
    Gloc = fdga.siam_bare_Green(w["nG"], w["T"], e=0.1, Δ=0.8, D=None)
    NP = w["L"] ** 2
    Gk = np.repeat(Gloc[:, None], NP, axis=1)
    S_nl = fdga.NL_ParquetSolver(w["nK1"], w["nK2"], w["nK3"], w["L"], Gk, np.zeros_like(Gk), np.zeros_like(Gk),
                                 T=w["T"], U=w["U"], mΠν_factor=w["mPinu"], precision=precision)
    S_loc = fdga.ParquetSolver(w["nK1"], w["nK2"], w["nK3"], Gloc, np.zeros_like(Gloc), np.zeros_like(Gloc), T=w["T"],
                               U=w["U"], mΠν_factor=w["mPinu"], precision=precision)
    # the same random P-independent starting vertex (amplitude 0.05 U, K3 damped at large ν, ν')
    rng = np.random.default_rng(seed)
    for ch in range(3):
        for cls in (1, 2, 3):
            n = S_loc.vertex(0, ch, cls).size
            v = 0.05 * w["U"] * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
            S_loc.set_vertex(0, ch, cls, v)
            S_nl.set_vertex(0, ch, cls, np.tile(v, NP))
    return S_nl, S_loc, NP


def _compare(S_nl, S_loc, NP, tol, what):
    for ch in range(3):
        for cls in (1, 2, 3):
            a = S_nl.vertex(0, ch, cls).reshape(-1, NP, order="F")
            b = S_loc.vertex(0, ch, cls)
            e = max(relerr(a[:, p], b) for p in (0, 1, 17, 136, NP - 1))
            assert e < tol, f"{what} ch={ch} cls={cls}: {e:.3e}"
    a = S_nl.Σ
    b = S_loc.Σ[:, 0]
    e = max(relerr(a[:, k], b) for k in range(a.shape[1]))
    assert e < tol, f"{what} Σ: {e:.3e}"


@pytest.mark.parametrize("strategy", ["fdPA", "scPA"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-5)])
@pytest.mark.parametrize("workload", ["L16", "L24"])
def test_lattice_solver_with_k_independent_data_is_the_local_solver(precision, tol, strategy, workload):
    import fdga
    S_nl, S_loc, NP = _pair(precision, W if workload == "L16" else W24)
    for it in range(2):
        fdga.iterate_solver(S_nl, strategy)
        fdga.iterate_solver(S_loc, strategy)
        _compare(S_nl, S_loc, NP, tol, f"{precision} {strategy} iteration {it}")
    # every momentum of the vertex is the same (the check above samples five; here all of K1)
    k1 = S_nl.vertex(0, 2, 1).reshape(-1, NP, order="F")
    assert relerr(k1, np.repeat(k1[:, :1], NP, axis=1)) < tol


C1 = dict(T=0.05, U=2.0, Delta=np.pi / 5, nK1=128, nK2=(64, 64), nK3=(32, 64), mPinu=6, nG=256)


@pytest.mark.parametrize("sym,iters", [(True, 3), (False, 2)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-5)])
def test_headline_workload_iterations_match_oracle(sym, iters, precision, tol):
    """The SIAM fdPA workload of the benchmark at its full boxes (nK1 = 128, nK2 = (64, 64), nK3 =
    (32, 64), mΠν = 6, nG = 256): iterations from the solver's initial state on the GPU (tile shapes
    from the autotuner) and on the oracle, with and without symmetry groups."""
    import fdga
    import pyoracle as po
    w = C1
    S = fdga.parquet_solver_siam_parquet_approximation(w["nG"], w["nK1"], w["nK2"], w["nK3"], e=0.0, Δ=w["Delta"],
                                                       D=None, T=w["T"], U=w["U"], mΠν_factor=w["mPinu"],
                                                       precision=precision)
    O = po.Oracle(nG=w["nG"], nK1=w["nK1"], nK2=w["nK2"], nK3=w["nK3"], mPinu=w["mPinu"], T=w["T"], U=w["U"])
    O.siam(0.0, w["Delta"], None)
    O.init()
    if sym:
        S.init_sym_grp()
        O.init_sym_grp()
    S.autotune(reps=1)
    for it in range(iters):
        fdga.iterate_solver(S, "fdPA")
        O.iterate(po.FDPA)
        e = relerr(fdga.flatten(S), O.flatten())
        assert e < tol, f"{precision} sym={sym} iteration {it}: {e:.3e}"


@pytest.mark.timeout(400)
def test_headline_workload_solved_to_1e_8_matches_oracle():
    """The SIAM fdPA workload at its full boxes, solved to max|R| < 1e-8 from the initial state with the
    reference's Anderson mixing (m = 8, β = 0.85) and the symmetry groups, on the GPU and on the oracle
    (OpenMP): the same number of iterations, the same fixed point (1e-9) and the same χ_sp, χ_ch."""
    import fdga
    import pyoracle as po
    w = C1
    S = fdga.parquet_solver_siam_parquet_approximation(w["nG"], w["nK1"], w["nK2"], w["nK3"], e=0.0, Δ=w["Delta"],
                                                       D=None, T=w["T"], U=w["U"], mΠν_factor=w["mPinu"])
    S.init_sym_grp()
    print("c1 solve: GPU", flush=True)
    it, res = fdga.solve(S, strategy="fdPA", tol=1e-8, maxiter=100)
    print(f"c1 solve: GPU {it} iterations, max|R| {res:.2e}; oracle", flush=True)
    O = po.Oracle(nG=w["nG"], nK1=w["nK1"], nK2=w["nK2"], nK3=w["nK3"], mPinu=w["mPinu"], T=w["T"], U=w["U"])
    O.siam(0.0, w["Delta"], None)
    O.init()
    O.init_sym_grp()
    it_o, res_o = O.solve(po.FDPA, maxiter=100, tol=1e-8)
    print(f"c1 solve: oracle {it_o} iterations, max|R| {res_o:.2e}", flush=True)
    assert res < 1e-8 and res_o < 1e-8
    assert it == it_o, (it, it_o)
    assert relerr(fdga.flatten(S), O.flatten()) < 1e-9
    sp, ch = fdga.susceptibilities(S)
    sp_o, ch_o = fdga.chi_from_k1(O.vertex(po.VTX_F, po.ACH, 1), O.vertex(po.VTX_F, po.TCH, 1), w["U"])
    assert relerr(sp, sp_o) < 1e-9 and relerr(ch, ch_o) < 1e-9
    S.close()

"""Symmetry-class evaluation on the GPU (init_sym_grp!, src/ParquetSolver.jl:203-231; lattice
solver src/nonlocal/ParquetSolver.jl:198-300).

With the symmetry groups installed, every SG-evaluated stage gives op(value of the class
representative) at the other members (the lattice groups add the square lattice's reflection and
rotation and the momentum forms of the frequency symmetries).  The inputs of
`make_pair` are random and NOT symmetric, so the comparison with the oracle checks the class
partition, the choice of representative and every member's operation exactly: a different
representative or a missing conjugation changes the result at O(1), not at rounding level.

The golden tests run the reference's own golden solves (test/test_siam_scPA.jl:28-31, 60-63;
test/test_siam_fdPA.jl:29-33, 84) end to end on the GPU path, with the symmetry groups the
reference tests install and the mΠν_factor = 1 bubble mesh the goldens were produced with.
"""
import numpy as np
import pytest

import pyoracle as po
from helpers import CASES, EXTRA_CASES, make_pair, relerr

pytestmark = pytest.mark.gpu
TOL = 1e-12
CH_ORDER = (po.PCH, po.ACH, po.TCH)
# nl_L16: 45 of 256 momenta hold K3 representatives (the operand builders skip the rest)
SYM_CASES = ["local_ref", "local_vertexref", "nl_L2", "nl_L3", "nl_LG4_L2", "nl_L6", "nl_L16"]
ALL_CASES = {**CASES, **EXTRA_CASES}


@pytest.fixture(scope="module")
def fdga_mod():
    import fdga
    return fdga


def _pair_sym(case, seed=0):
    S, O = make_pair(case, seed=seed)
    S.init_sym_grp()
    O.init_sym_grp()
    return S, O


def _cmp_vertex(S, O, vtx_h, vtx_o, classes, what):
    from fdga import _lib as L
    for ch in range(3):
        for cls in classes:
            e = relerr(S.get(L.vertex_id(vtx_h, ch, cls)), O.get(po.vertex_id(vtx_o, ch, cls)))
            assert e < TOL, f"{what} ch={ch} cls={cls}: rel err {e:.3e}"


@pytest.mark.parametrize("case", SYM_CASES)
def test_sym_bse_stages(case, fdga_mod):
    S, O = _pair_sym(case)
    fdga_mod.build_K3_cache(S)
    O.build_k3_cache()
    for ch in CH_ORDER:
        fdga_mod.BSE_L_K2(S, ch)
        O.bse_l_k2(ch)
    _cmp_vertex(S, O, 2, po.VTX_FL, (2,), f"{case} L_K2")
    for ch in CH_ORDER:
        fdga_mod.BSE_L_K3(S, ch)
        O.bse_l_k3(ch)
    _cmp_vertex(S, O, 2, po.VTX_FL, (3,), f"{case} L_K3")
    for ch in CH_ORDER:
        fdga_mod.BSE_K1(S, ch)
        O.bse_k1(ch)
    _cmp_vertex(S, O, 1, po.VTX_FBUFF, (1,), f"{case} K1")
    for ch in CH_ORDER:
        fdga_mod.BSE_K2(S, ch)
        O.bse_k2(ch)
    _cmp_vertex(S, O, 1, po.VTX_FBUFF, (2,), f"{case} K2")
    for ch in CH_ORDER:
        fdga_mod.BSE_K3(S, ch)
        O.bse_k3(ch)
    _cmp_vertex(S, O, 1, po.VTX_FBUFF, (3,), f"{case} K3")


@pytest.mark.parametrize("case", SYM_CASES)
@pytest.mark.parametrize("strategy", ["scPA", "fdPA"])
def test_sym_sde(case, strategy, fdga_mod):
    from fdga import _lib as L
    S, O = _pair_sym(case)
    fdga_mod.SDE(S, strategy)
    O.sde(po.SCPA if strategy == "scPA" else po.FDPA)
    arrays = [(L.SIGMA, po.SIGMA, "Σ")]
    # Lpp/Lph hold the last SDE level (the RefVertex); with null tables that level is exactly
    # zero, the GPU path skips it and its Lpp/Lph keep the previous level
    if ALL_CASES[case][4]:
        arrays += [(L.LPP, po.LPP, "Lpp"), (L.LPH, po.LPH, "Lph")]
    for hw, ow, name in arrays:
        e = relerr(S.get(hw), O.get(ow))
        assert e < TOL, f"{case} {strategy} {name}: rel err {e:.3e}"


@pytest.mark.parametrize("case", SYM_CASES)
@pytest.mark.parametrize("strategy", ["scPA", "fdPA"])
def test_sym_full_iteration(case, strategy, fdga_mod):
    S, O = _pair_sym(case, seed=3)
    st = po.SCPA if strategy == "scPA" else po.FDPA
    for it in range(2):
        fdga_mod.iterate_solver(S, strategy)
        O.iterate(st)
        e = relerr(fdga_mod.flatten(S), O.flatten())
        assert e < TOL, f"{case} {strategy} iteration {it}: {e:.3e}"


def test_sym_changes_result_and_reset_restores_dense(fdga_mod):
    """On asymmetric inputs the class evaluation differs from the dense one at O(1) (so the
    tests above are sensitive to it); reset_sym_grp! gives back the dense scPA result."""
    S, O = make_pair("local_ref", seed=11)
    x0 = fdga_mod.flatten(S)
    S.init_sym_grp()
    fdga_mod.iterate_solver(S, "scPA")
    x_sym = fdga_mod.flatten(S)
    fdga_mod.unflatten(S, x0)
    S.reset_sym_grp()
    fdga_mod.iterate_solver(S, "scPA")
    O.iterate(po.SCPA)
    x_dense = fdga_mod.flatten(S)
    assert relerr(x_dense, O.flatten()) < TOL
    assert relerr(x_sym, x_dense) > 1e-3


def test_sym_lattice_skips_tiles_and_changes_result(fdga_mod):
    """Lattice groups: the K3 contractions compute only the tiles holding representatives, the
    class evaluation differs from the dense one at O(1) on asymmetric inputs, and reset_sym_grp!
    gives back the dense result."""
    S, O = make_pair("nl_L6", seed=11)
    x0 = fdga_mod.flatten(S)
    dense_flops = [S.gemm_flops(left) for left in (1, 0)]
    S.init_sym_grp()
    sym_flops = [S.gemm_flops(left) for left in (1, 0)]
    assert all(s < 0.5 * d for s, d in zip(sym_flops, dense_flops)), (sym_flops, dense_flops)
    # scPA: reset_sym_grp! leaves SG0pp2 / SG0ph2 (fdPA's reference system) installed
    fdga_mod.iterate_solver(S, "scPA")
    x_sym = fdga_mod.flatten(S)
    fdga_mod.unflatten(S, x0)
    S.reset_sym_grp()
    fdga_mod.iterate_solver(S, "scPA")
    O.iterate(po.SCPA)
    x_dense = fdga_mod.flatten(S)
    assert relerr(x_dense, O.flatten()) < TOL
    assert relerr(x_sym, x_dense) > 1e-3


# ----------------------------------------------------------------------- reference goldens
T, U = 0.1, 1.0
# `S.F.γa.K1.(2π * T .* -2:2)` parses as the range (-4πT):1:2: 4 points, 3 of them off-mesh,
# evaluated as linear interpolation on the bosonic mesh
XS = np.arange(2 * np.pi * T * -2, 2 + 1e-12, 1.0)


def _interp(K, nK1):
    vals = 2 * np.pi * T * np.arange(-(nK1 - 1), nK1)
    return np.interp(XS, vals, K.real) + 1j * np.interp(XS, vals, K.imag)


def _scpa(fdga_mod, e, D, nmax, nK2, nK3, Delta=np.pi / 5):
    nG, nK1 = 6 * nmax, 4 * nmax
    S = fdga_mod.parquet_solver_siam_parquet_approximation(nG, nK1, nK2, nK3, e=e, Δ=Delta, D=D, T=T, U=U,
                                                           mΠν_factor=1)
    S.init_sym_grp()
    return S, nG, nK1


def test_golden_siam_scpa_halffilling(fdga_mod):
    """test/test_siam_scPA.jl:19-31 on the GPU path (NLsolve-style Anderson, ftol = 1e-4)."""
    S, nG, nK1 = _scpa(fdga_mod, 0.0, 10.0, 6, (6, 6), (6, 6))
    it, res = fdga_mod.solve(S, strategy="scPA", maxiter=100, tol=1e-4)
    assert res < 1e-4
    gS = [-0.052138235296134906, -0.03838544776344314, 0.03838544776344314, 0.052138235296134906]
    np.testing.assert_allclose(S.Σ[[n + nG for n in (-2, -1, 0, 1)], 0], gS, rtol=3e-4)
    gold = {
        po.ACH: [0.13203850929270397, 0.5403615530152339, 0.2333246221064017, 0.09056300899983459],
        po.PCH: [-0.10420799999591804, -0.2403951910434166, -0.15592452265704748, -0.07622568434721624],
        po.TCH: [0.013898648018808482, 0.1499562726081748, 0.03867632419082161, 0.007160400841240708],
    }
    for ch, g in gold.items():
        np.testing.assert_allclose(_interp(S.vertex(0, ch, 1), nK1), g, rtol=3e-4)
    # the symmetrized half-filled solution: iΣ real and odd in ν
    Sall = S.Σ[:, 0]
    assert np.max(np.abs(Sall.imag)) < 1e-12
    np.testing.assert_allclose(Sall, -Sall[::-1], atol=1e-12)


def test_golden_siam_scpa_doped(fdga_mod):
    """test/test_siam_scPA.jl:46-63 on the GPU path."""
    S, nG, nK1 = _scpa(fdga_mod, 0.5, 10.0, 6, (7, 6), (7, 6))
    fdga_mod.solve(S, strategy="scPA", maxiter=100, tol=1e-4)
    gS = [-0.0389123277075552 - 0.16855090184215607j, -0.025252640312580586 - 0.17429637478745583j,
          0.025252640312580586 - 0.17429637478745583j, 0.0389123277075552 - 0.16855090184215607j]
    np.testing.assert_allclose(S.Σ[[n + nG for n in (-2, -1, 0, 1)], 0], gS, rtol=1e-4)
    gold = {
        po.ACH: [0.11925962005661812 + 8.57514999021054e-5j, 0.416232811242488 + 3.319936929625957e-5j,
                 0.20353141073439696 - 8.209974062547027e-5j, 0.08259294412067451 - 7.660306021755952e-5j],
        po.PCH: [-0.12570450372739117 + 0.06583917638195431j, -0.24548654160724023 + 0.021014183409764874j,
                 -0.17578586892780296 - 0.05408344507941078j, -0.09544981624337806 - 0.06686768343644132j],
        po.TCH: [0.011016969129875598 + 7.455601769977568e-5j, 0.09533547032272821 + 2.4477973720973318e-5j,
                 0.028843799251846686 - 6.260706942592786e-5j, 0.005828299701446311 - 7.32415771550901e-5j],
    }
    for ch, g in gold.items():
        np.testing.assert_allclose(_interp(S.vertex(0, ch, 1), nK1), g, atol=3e-4 * np.max(np.abs(g)))


def test_siam_sym_solve_matches_oracle(fdga_mod):
    """The symmetrized scPA solve to 1e-12 agrees with the oracle's symmetrized solve to 1e-10."""
    S, nG, nK1 = _scpa(fdga_mod, 0.0, 10.0, 6, (6, 6), (6, 6))
    it, res = fdga_mod.solve(S, strategy="scPA", tol=1e-12, maxiter=80)
    O = po.Oracle(nG=nG, nK1=nK1, nK2=(6, 6), nK3=(6, 6), mPinu=1, T=T, U=U)
    O.siam(0.0, np.pi / 5, 10.0)
    O.init()
    O.init_sym_grp()
    it2, res2 = O.solve(po.SCPA, tol=1e-12, maxiter=80)
    assert res < 1e-12 and res2 < 1e-12
    assert relerr(fdga_mod.flatten(S), O.flatten()) < 1e-10


def _fdpa_pair(fdga_mod, nmax=6):
    D, e, Delta = 10.0, -0.3, np.pi / 3
    nG, nK1 = 8 * nmax, 8 * nmax
    out = []
    for _ in range(2):
        S = fdga_mod.parquet_solver_siam_parquet_approximation(nG, nK1, (nmax, nmax), (nmax, nmax), e=e, Δ=Delta, D=D,
                                                               T=T, U=U, mΠν_factor=1, fdpa_h0=False)
        S.init_sym_grp()
        out.append(S)
    return out


def test_golden_fdpa_trivial_reference_is_scpa(fdga_mod):
    """test/test_siam_fdPA.jl:19-33: fdPA from G0 = Σ0 = 0, F0 = U equals scPA to 1e-10."""
    a, b = _fdpa_pair(fdga_mod)
    fdga_mod.solve(a, strategy="scPA")
    fdga_mod.solve(b, strategy="fdPA")
    assert np.max(np.abs(a.Σ - b.Σ)) < 1e-10
    assert np.max(np.abs(fdga_mod.flatten(a) - fdga_mod.flatten(b))) < 1e-10


def test_golden_fdpa_box8(fdga_mod):
    """test/test_siam_fdPA.jl:36-84: scPA reference solve at (e=-0.3, Δ=π/3, D=10) with box 12,
    then fdPA with box 8 towards (e=0.5, Δ=π/5, D=20); golden Σ(πT)."""
    nref, nfd = 12, 8
    nG, nK1 = 8 * nref, 8 * nref
    s0 = fdga_mod.parquet_solver_siam_parquet_approximation(nG, nK1, (nref, nref), (nref, nref), e=-0.3, Δ=np.pi / 3,
                                                            D=10.0, T=T, U=U, mΠν_factor=1, fdpa_h0=False)
    s0.init_sym_grp()
    fdga_mod.solve(s0, strategy="scPA")
    F0 = {(ch, cls): s0.vertex(0, ch, cls) for ch in range(3) for cls in (1, 2, 3)}
    Gbare = fdga_mod.siam_bare_Green(nG, T, e=0.5, Δ=np.pi / 5, D=20.0)
    from fdga import _lib as L
    S = fdga_mod.ParquetSolver(8 * nfd, (nfd, nfd), (nfd, nfd), Gbare, s0.get(L.G), s0.get(L.SIGMA), F0, T=T, U=U,
                               mΠν_factor=1, fdpa_h0=False, F0_boxes=(nK1, (nref, nref), (nref, nref)))
    S.init_sym_grp()
    fdga_mod.solve(S, strategy="fdPA")
    np.testing.assert_allclose(S.Σ[nG, 0], 0.024643001835742997 - 0.17494219707558506j, rtol=2e-4)

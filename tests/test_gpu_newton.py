"""Newton–Krylov solve of the parquet fixed point and dχ/dU by the implicit function theorem, on
the GPU path (SIAM, the golden-test parameters of test/test_siam_scPA.jl:19-31 with symmetry
groups and the mΠν_factor = 1 bubble mesh)."""
import numpy as np
import pytest

from helpers import relerr

pytestmark = pytest.mark.gpu
T, NMAX = 0.1, 6


@pytest.fixture(scope="module")
def fdga_mod():
    import fdga
    return fdga


def _mk(fdga_mod, U):
    S = fdga_mod.parquet_solver_siam_parquet_approximation(6 * NMAX, 4 * NMAX, (NMAX, NMAX), (NMAX, NMAX), e=0.0,
                                                           Δ=np.pi / 5, D=10.0, T=T, U=U, mΠν_factor=1)
    S.init_sym_grp()
    return S


def test_newton_krylov_matches_mixing_with_fewer_iterations(fdga_mod):
    from fdga import newton
    A = _mk(fdga_mod, 1.0)
    it_a, res_a = fdga_mod.solve(A, strategy="scPA", tol=1e-10, maxiter=100)
    N = _mk(fdga_mod, 1.0)
    it_n, res_n, nev = newton.solve_newton(N, strategy="scPA", tol=1e-10)
    assert res_a < 1e-10 and res_n < 1e-10
    assert relerr(fdga_mod.flatten(N), fdga_mod.flatten(A)) < 1e-8
    assert it_n < it_a, (it_n, it_a)


def test_dchi_dU_implicit_function_theorem_vs_finite_differences(fdga_mod):
    """dχ_sp(Ω=0)/dU with χ_sp = K1a/U² from dx*/dU = -J_R^{-1} dR/dU, against central differences
    of full solves at U ± 1e-4."""
    from fdga import newton
    from fdga import _lib as L
    U, h = 1.0, 1e-4
    S = _mk(fdga_mod, U)
    fdga_mod.solve(S, strategy="scPA", tol=1e-12, maxiter=100)
    xs = fdga_mod.flatten(S)
    Sp, Sm = _mk(fdga_mod, U + h), _mk(fdga_mod, U - h)
    dR = (newton.solver_residual(Sp, "scPA")(xs) - newton.solver_residual(Sm, "scPA")(xs)) / (2 * h)
    dx, rel, _ = newton.fixed_point_sensitivity(newton.solver_residual(S, "scPA"), xs, dR, tol=1e-10)
    seg = newton.flat_segments(S)[(L.ACH, 1)]
    i0 = 4 * NMAX - 1                                   # Ω = 0 on the K1 mesh
    k1a, dk1a = xs[seg][i0], dx[seg][i0]
    dchi = dk1a / U**2 - 2 * k1a / U**3

    def chi(Uv):
        Sv = _mk(fdga_mod, Uv)
        fdga_mod.solve(Sv, strategy="scPA", tol=1e-12, maxiter=100)
        return Sv.vertex(0, L.ACH, 1)[i0] / Uv**2
    fd = (chi(U + h) - chi(U - h)) / (2 * h)
    assert abs(dchi - fd) < 1e-5 * abs(fd), (dchi, fd)

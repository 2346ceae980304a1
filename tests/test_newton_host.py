"""Host-side Newton–Krylov machinery (fdga.newton) on problems with known answers (no GPU)."""
import numpy as np

from fdga import newton


def _rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_gmres_matches_direct_solve():
    rng = np.random.default_rng(1)
    n = 60
    A = np.eye(n) * 4 + 0.3 * _rand_c(rng, n, n) / np.sqrt(n)
    b = _rand_c(rng, n)
    x, rel, nprod = newton.gmres(lambda v: A @ v, b, tol=1e-12, restart=40, maxiter=5)
    assert rel < 1e-12
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
    assert nprod <= 80


def test_gmres_restarts_converge():
    rng = np.random.default_rng(2)
    n = 80
    A = np.eye(n) * 2 + 0.6 * _rand_c(rng, n, n) / np.sqrt(n)
    b = _rand_c(rng, n)
    x, rel, _ = newton.gmres(lambda v: A @ v, b, tol=1e-10, restart=10, maxiter=40)
    assert rel < 1e-10
    np.testing.assert_allclose(A @ x, b, atol=1e-8)


def _toy(rng, n=40, p=1.0):
    """x = G_p(x) = p (A x + 0.05 x^2) + c: a contraction with a unique fixed point near c."""
    A = 0.5 * _rand_c(rng, n, n) / np.sqrt(n)
    c = 0.3 * _rand_c(rng, n)

    def G(x, p=p):
        return p * (A @ x + 0.05 * x * x) + c
    return A, c, G


def test_newton_krylov_toy_fixed_point():
    rng = np.random.default_rng(3)
    A, c, G = _toy(rng)
    x, it, res, nev = newton.newton_krylov(lambda x: G(x) - x, np.zeros_like(c), tol=1e-12, maxiter=20)
    assert res < 1e-12
    np.testing.assert_allclose(G(x), x, atol=1e-12)
    # Newton: a handful of outer iterations (plain iteration of G needs dozens)
    assert it <= 6


def test_fixed_point_sensitivity_linear_closed_form():
    """G_p(x) = p A x + c: x* = (I - pA)^{-1} c and dx*/dp = (I - pA)^{-1} A x*."""
    rng = np.random.default_rng(4)
    n = 30
    A = 0.4 * _rand_c(rng, n, n) / np.sqrt(n)
    c = _rand_c(rng, n)
    p = 0.8
    M = np.eye(n) - p * A
    xs = np.linalg.solve(M, c)
    R = lambda x: p * (A @ x) + c - x  # noqa: E731
    h = 1e-5
    dR_dp = ((p + h) * (A @ xs) - (p - h) * (A @ xs)) / (2 * h)
    dx, rel, _ = newton.fixed_point_sensitivity(R, xs, dR_dp, tol=1e-12)
    np.testing.assert_allclose(dx, np.linalg.solve(M, A @ xs), rtol=1e-6, atol=1e-8)


def test_fixed_point_sensitivity_nonlinear_against_finite_differences():
    rng = np.random.default_rng(5)
    A, c, G = _toy(rng)
    p0, h = 1.0, 1e-5

    def solve(p):
        x, _, res, _ = newton.newton_krylov(lambda x: G(x, p) - x, np.zeros_like(c), tol=1e-13, maxiter=30)
        assert res < 1e-13
        return x
    xs = solve(p0)
    dR_dp = (G(xs, p0 + h) - G(xs, p0 - h)) / (2 * h)
    dx, _, _ = newton.fixed_point_sensitivity(lambda x: G(x, p0) - x, xs, dR_dp, tol=1e-12)
    fd = (solve(p0 + h) - solve(p0 - h)) / (2 * h)
    np.testing.assert_allclose(dx, fd, rtol=1e-5, atol=1e-7)


def test_newton_stops_when_no_step_lowers_the_residual():
    """R(x) = x with the linear solves made to return the step of the wrong sign: no trial of the
    backtracking lowers ||R||, so the iteration keeps the last accepted x and reports a residual above
    tol instead of taking a worsening step."""
    x0 = np.full(4, 1.0 + 0.5j)
    orig = newton.gmres

    def gmres_wrong_sign(matvec, b, **kw):
        x, rel, n = orig(matvec, b, **kw)
        return -x, rel, n
    newton.gmres = gmres_wrong_sign
    try:
        x, it, res, nev = newton.newton_krylov(lambda x: x.copy(), x0, tol=1e-10, maxiter=5)
    finally:
        newton.gmres = orig
    assert res >= 1e-10
    np.testing.assert_array_equal(x, x0)

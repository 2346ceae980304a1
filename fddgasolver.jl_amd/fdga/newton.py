"""Newton–Krylov solve of the parquet fixed point and fixed-point sensitivities.

The reference solves the fixed point x = G(x) (`fixed_point!`, src/solve.jl:119-157) with NLsolve's
Anderson mixing (`solve!`, src/solve.jl:160-196), or preconditions it with its linearised BSE
(`mfRGLinearMap`, src/mfRG.jl:20-89, used by `fixed_point_preconditioned!`, src/mfRG.jl:93-173).
This module adds the Newton alternative: Newton's method on R(x) = G(x) - x with the Jacobian
applied matrix-free, J_R v = (R(x + h v) - R(x)) / h, each product one parquet iteration on the GPU,
and the linear systems solved by restarted GMRES.  The same machinery gives the derivative of the
fixed point with respect to a model parameter p (implicit function theorem):
    dx*/dp = -J_R(x*)^{-1} dR/dp,   dR/dp = (G_{p+h}(x*) - G_{p-h}(x*)) / (2h),
from which dχ/dU follows without re-solving the fixed point at U ± h.

The functions take plain callables (x -> R(x)) so that they run on any residual; `solver_residual`
wraps a device solver (`fixed_point`, src/solve.jl:119-157).
"""
import numpy as np


def _norm(v):
    return float(np.linalg.norm(v))


def gmres(matvec, b, *, tol=1e-6, restart=30, maxiter=5, x0=None):
    """Restarted GMRES (modified Gram–Schmidt Arnoldi, least-squares on the Hessenberg matrix) for
    the complex system A x = b given x -> A x.  Stops when ||b - A x|| <= tol ||b||.  Returns
    (x, relative residual, number of products)."""
    b = np.asarray(b)
    bn = _norm(b)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=b.dtype)
    if bn == 0.0:
        return x, 0.0, 0
    nprod = 0
    rel = 1.0
    for _ in range(maxiter):
        r = b - matvec(x) if x0 is not None or nprod > 0 else b.copy()
        if x0 is not None or nprod > 0:
            nprod += 1
        beta = _norm(r)
        rel = beta / bn
        if rel <= tol:
            break
        V = [r / beta]
        H = np.zeros((restart + 1, restart), dtype=np.complex128)
        y = None
        for j in range(restart):
            w = matvec(V[j])
            nprod += 1
            for i in range(j + 1):
                H[i, j] = np.vdot(V[i], w)
                w = w - H[i, j] * V[i]
            H[j + 1, j] = _norm(w)
            e1 = np.zeros(j + 2, dtype=np.complex128)
            e1[0] = beta
            y = np.linalg.lstsq(H[: j + 2, : j + 1], e1, rcond=None)[0]
            res = _norm(e1 - H[: j + 2, : j + 1] @ y)
            rel = res / bn
            if rel <= tol or H[j + 1, j] == 0.0:
                break
            V.append(w / H[j + 1, j])
        x = x + sum(y[i] * V[i] for i in range(len(y)))
        if rel <= tol:
            break
    return x, rel, nprod


def newton_krylov(residual, x0, *, tol=1e-8, maxiter=30, eta=0.1, restart=30, gmres_restarts=3, h_rel=1e-7,
                  verbose=False):
    """Newton's method on R(x) = 0 with matrix-free Jacobian products and GMRES.

    The Newton step solves J dx = -R to the relative tolerance eta (Eisenstat–Walker choice 2,
    bounded by `eta`); J v = (R(x + h v) - R(x)) / h with h = h_rel (1 + ||x||_inf) / ||v||_inf.
    The step is accepted when it lowers ||R||_2, else halved (up to 8 times); when no trial lowers it the
    iteration stops at the last accepted x (max|R| >= tol tells the caller it did not converge).
    Converged when max|R| < tol, as in `solve!`.  Returns (x, newton iterations, max|R|, residual
    evaluations)."""
    x = np.array(x0, dtype=np.complex128)
    R = residual(x)
    nev = 1
    res = float(np.max(np.abs(R)))
    rn_old = None
    it = 0
    for it in range(1, maxiter + 1):
        if res < tol:
            it -= 1
            break
        rn = _norm(R)
        # forcing term: Eisenstat–Walker 2 (gamma = 0.9, alpha = 2), capped at eta
        eta_k = eta if rn_old is None else min(eta, 0.9 * (rn / rn_old) ** 2)
        eta_k = max(eta_k, 1e-4)
        xs = 1.0 + float(np.max(np.abs(x)))

        def jv(v):
            vn = float(np.max(np.abs(v)))
            if vn == 0.0:
                return np.zeros_like(v)
            h = h_rel * xs / vn
            return (residual(x + h * v) - R) / h

        dx, lin_rel, nprod = gmres(jv, -R, tol=eta_k, restart=restart, maxiter=gmres_restarts)
        nev += nprod
        lam = 1.0
        accepted = False
        for _ in range(9):
            xn = x + lam * dx
            Rn = residual(xn)
            nev += 1
            if _norm(Rn) < rn:
                accepted = True
                break
            lam *= 0.5
        if not accepted:
            if verbose:
                print(f"newton {it:3d}  no step lowers ||R||: stopping at max|R| = {res:.3e}")
            break
        x, R, rn_old = xn, Rn, rn
        res = float(np.max(np.abs(R)))
        if verbose:
            print(f"newton {it:3d}  max|R| = {res:.3e}  gmres products {nprod}  rel {lin_rel:.1e}  step {lam}")
    return x, it, res, nev


def fixed_point_sensitivity(residual, x_star, dR_dp, *, tol=1e-8, restart=40, gmres_restarts=4, h_rel=1e-7):
    """dx*/dp = -J_R(x*)^{-1} dR/dp at a fixed point x* of G, with J_R applied matrix-free.
    Returns (dx*/dp, GMRES relative residual, residual evaluations)."""
    x = np.asarray(x_star, dtype=np.complex128)
    R0 = residual(x)
    xs = 1.0 + float(np.max(np.abs(x)))

    def jv(v):
        vn = float(np.max(np.abs(v)))
        if vn == 0.0:
            return np.zeros_like(v)
        h = h_rel * xs / vn
        return (residual(x + h * v) - R0) / h

    dx, rel, nprod = gmres(jv, -np.asarray(dR_dp), tol=tol, restart=restart, maxiter=gmres_restarts)
    return dx, rel, nprod + 1


def solver_residual(S, strategy="fdPA"):
    """x -> R(x) = G(x) - x of a device solver (fixed_point!, src/solve.jl:119-157)."""
    from .solver import fixed_point

    def R(x):
        out = np.empty_like(np.asarray(x, dtype=np.complex128))
        fixed_point(out, x, S, strategy)
        return out
    return R


def flat_segments(S):
    """Offsets of the pieces of the flattened state [γp(K1,K2,K3); γt(...); γa(...); Σ]
    (src/ParquetSolver.jl:277-306): dict (ch, cls) -> slice, and 'Sigma'."""
    from . import _lib as L
    n1 = S.size(L.vertex_id(0, 0, 1))
    n2 = S.size(L.vertex_id(0, 0, 2))
    n3 = S.size(L.vertex_id(0, 0, 3))
    seg, off = {}, 0
    for ch in (L.PCH, L.TCH, L.ACH):
        for cls, n in ((1, n1), (2, n2), (3, n3)):
            seg[(ch, cls)] = slice(off, off + n)
            off += n
    seg["Sigma"] = slice(off, off + S.size(L.SIGMA))
    return seg


def solve_newton(S, *, strategy="fdPA", tol=1e-8, maxiter=30, eta=0.1, restart=30, verbose=False):
    """Newton–Krylov counterpart of `solve!` (src/solve.jl:160-196): the device solver S ends at the
    fixed point.  Returns (newton iterations, max|R|, parquet iterations used)."""
    from .solver import flatten, unflatten
    x, it, res, nev = newton_krylov(solver_residual(S, strategy), flatten(S), tol=tol, maxiter=maxiter, eta=eta,
                                    restart=restart, verbose=verbose)
    unflatten(S, x)
    return it, res, nev

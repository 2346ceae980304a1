"""The host Anderson loop of fdga.solve (solve!, src/solve.jl:160-196, NLsolve's :anderson) on a
linear contraction G(x) = A x + b standing in for the device iteration.  The fixed point solves
(I - A) x = b; without truncation (mem >= n) Anderson acceleration on a linear map is essentially
GMRES (Walker & Ni, SIAM J. Numer. Anal. 49 (2011) 1715), so it ends within n + 2 evaluations,
while the plain damped iteration x + δ f (mem = 0) needs many more."""
import numpy as np
import pytest

from fdga import solver

N = 12


def _problem(seed=2):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    A *= 0.9 / np.max(np.abs(np.linalg.eigvals(A)))
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    return A, b, np.linalg.solve(np.eye(N) - A, b)


def _run(monkeypatch, mem, maxiter=400):
    A, b, xs = _problem()
    last = {}

    def fixed_point(R, x, S, strategy, update_Σ=True, compute_Hartree=True):
        last["x"] = x.copy()
        R[:] = A @ x + b - x
        return R

    monkeypatch.setattr(solver, "flatten", lambda S: np.zeros(N, dtype=np.complex128))
    monkeypatch.setattr(solver, "fixed_point", fixed_point)
    it, res = solver.solve(None, strategy="fdPA", maxiter=maxiter, tol=1e-10, mem=mem)
    return it, res, float(np.max(np.abs(last["x"] - xs)))


def test_untruncated_anderson_ends_like_gmres(monkeypatch):
    it, res, err = _run(monkeypatch, mem=N)
    assert res < 1e-10 and err < 1e-9
    assert it <= N + 2, it


@pytest.mark.parametrize("mem", [4, 0])
def test_truncated_and_plain_mixing_reach_the_same_fixed_point(monkeypatch, mem):
    it_full = _run(monkeypatch, mem=N)[0]
    it, res, err = _run(monkeypatch, mem=mem)
    assert res < 1e-10 and err < 1e-8
    assert it > it_full

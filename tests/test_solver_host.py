"""Argument checks of the host solver API that run before any device call, with the reference's
messages: the Channel constructor's box assertions (src/channel.jl:33-34,40) and iterate_solver!'s
strategy assertion (src/solve.jl:10); the reference strategy whose BSE form is not on the device
path (the one-loop K3 of :fdPA_1loop) is refused by name; the state vectors handed to the C side are
checked for length, dtype and contiguity before the call."""
import numpy as np
import pytest

import fdga
from fdga import _lib as L
from fdga import solver


@pytest.mark.parametrize("nK1,nK2,nK3,msg", [
    (3, (4, 2), (1, 1), "bosonic frequencies in K2"),
    (3, (2, 4), (1, 1), "fermionic frequencies in K2"),
    (4, (2, 2), (3, 1), "larger than in K3"),
    (4, (2, 2), (1, 3), "larger than in K3"),
])
def test_box_assertions(nK1, nK2, nK3, msg):
    with pytest.raises(AssertionError, match=msg):
        fdga.Solver(nG=8, nK1=nK1, nK2=nK2, nK3=nK3, T=0.1, U=1.0, mPinu=1)


def test_strategy_names():
    assert solver._strategy("scPA") == L.SCPA and solver._strategy("fdPA") == L.FDPA
    assert solver._strategy("scPA_new") == L.SCPA_NEW and solver._strategy("fdPA_new") == L.FDPA_NEW
    with pytest.raises(AssertionError, match="Calculation strategy unknown"):
        solver._strategy("mfRG")
    with pytest.raises(NotImplementedError, match="fdPA_1loop"):
        solver._strategy("fdPA_1loop")


class _Lib:
    """stand-in for libfdga.so: sizes only; reaching a compute entry point fails the test"""
    def fdga_flat_size(self, h):
        return 10

    def fdga_flat_size_F(self, h):
        return 7

    def __getattr__(self, name):
        raise AssertionError(f"{name} reached with an unchecked vector")


class _S:
    lib, h = _Lib(), None


@pytest.mark.parametrize("update", [True, False])
def test_fixed_point_checks_vectors(update):
    n = 10 if update else 7
    good = np.zeros(n, complex)
    with pytest.raises(ValueError, match="x"):
        solver.fixed_point(np.zeros(n, complex), np.zeros(n + 1, complex), _S(), "fdPA", update)
    with pytest.raises(ValueError, match="R"):
        solver.fixed_point(np.zeros(n - 1, complex), good, _S(), "fdPA", update)
    with pytest.raises(ValueError, match="R"):
        solver.fixed_point(np.zeros(n, np.float64), good, _S(), "fdPA", update)
    with pytest.raises(ValueError, match="R"):
        solver.fixed_point(np.zeros(2 * n, complex)[::2], good, _S(), "fdPA", update)


def test_unflatten_checks_length():
    with pytest.raises(ValueError, match="unflatten"):
        solver.unflatten(_S(), np.zeros(9, complex))


def test_mfrg_map_checks_length_and_strategy():
    with pytest.raises(ValueError, match="Invalid strategy"):
        solver.mfRGLinearMap(_S(), "scPA")
    A = solver.mfRGLinearMap(_S(), "fdPA")
    assert A.shape == (7, 7)
    with pytest.raises(ValueError, match="mfRGLinearMap"):
        A.matvec(np.zeros(8, complex))

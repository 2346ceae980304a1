"""Argument checks of the host solver API that run before any device call, with the reference's
messages: the Channel constructor's box assertions (src/channel.jl:33-34,40) and iterate_solver!'s
strategy assertion (src/solve.jl:10); the reference strategies whose BSE forms are not on the
device path (the alternative K1/K2 forms, the one-loop K3) are refused by name."""
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
    with pytest.raises(AssertionError, match="Calculation strategy unknown"):
        solver._strategy("mfRG")
    for s in ("scPA_new", "fdPA_new", "fdPA_1loop"):
        with pytest.raises(NotImplementedError, match=s):
            solver._strategy(s)

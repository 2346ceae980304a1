"""χ read-out of a device solver (fdga.susceptibilities: K1a and K1t read back through the C-ABI)
against the same formula applied to the oracle's K1 after two fdPA iterations from identical
random inputs, for the local and the s-wave lattice solver, FP64 (1e-12 relative, as the other
FP64 parity tests)."""
import pytest

import pyoracle as po
from helpers import make_pair, relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["local_vertexref", "nl_L3"])
def test_susceptibilities_match_oracle(case):
    import fdga
    S, O = make_pair(case, seed=7)
    for _ in range(2):
        fdga.iterate_solver(S, "fdPA")
        O.iterate(po.FDPA)
    sp, ch = fdga.susceptibilities(S)
    NP = S.NP if S.nonlocal_ else 1
    osp, och = fdga.chi_from_k1(O.vertex(po.VTX_F, po.ACH, 1), O.vertex(po.VTX_F, po.TCH, 1), S.U, NP)
    assert sp.shape == osp.shape and ch.shape == och.shape
    e = max(relerr(sp, osp), relerr(ch, och))
    assert e < 1e-12, f"{case}: rel err {e:.3e}"

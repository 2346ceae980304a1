"""χ read-out (src/utility/load_triqs.jl:73-75,106-115) on the host: the pSp-storage formula
χ_sp = K1a/U², χ_ch = -(2 K1t - K1a)/U² checked against the density/magnetic basis it comes from
(K1_D = -U² χ_D, K1_M = -U² χ_M, K1a = -K1_M, K1t = (K1_D - K1_M)/2), so that exchanging the a and t
channels or dropping a term fails; and the lattice layout (Ω fastest, then P) through a stand-in
solver that serves K1 arrays the way the device read-back does."""
import numpy as np
import pytest

import fdga


def _dm_pair(rng, shape):
    χD = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    χM = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return χD, χM


def _k1_from_dm(χD, χM, U):
    K1D, K1M = -U * U * χD, -U * U * χM
    return -K1M, 0.5 * (K1D - K1M)   # K1a, K1t


@pytest.mark.parametrize("U", [0.5, 2.0, 3.7])
def test_chi_from_k1_recovers_the_density_magnetic_basis(U):
    rng = np.random.default_rng(11)
    χD, χM = _dm_pair(rng, 2 * 17 - 1)
    k1a, k1t = _k1_from_dm(χD, χM, U)
    sp, ch = fdga.chi_from_k1(k1a, k1t, U)
    np.testing.assert_allclose(sp, χM, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(ch, χD, rtol=1e-14, atol=1e-14)


def test_chi_from_k1_rejects_mismatched_arrays():
    with pytest.raises(ValueError):
        fdga.chi_from_k1(np.zeros(5), np.zeros(7), 2.0)


class _StandIn:
    """Serves K1 of the full vertex (vertex 0) as the solver's read-back does: flat, Ω fastest."""

    def __init__(self, k1a, k1t, U, NP):
        self.U, self.NP, self.nonlocal_ = U, NP, NP > 1
        self._k1 = {fdga.aCh: k1a.reshape(-1, order="F"), fdga.tCh: k1t.reshape(-1, order="F")}

    def vertex(self, vtx, ch, cls):
        assert vtx == 0 and cls == 1
        return self._k1[ch].copy()


@pytest.mark.parametrize("NP", [1, 16])
def test_susceptibilities_layout(NP):
    U, nK1 = 2.0, 9
    rng = np.random.default_rng(5)
    shape = (2 * nK1 - 1, NP) if NP > 1 else (2 * nK1 - 1,)
    χD, χM = _dm_pair(rng, shape)
    k1a, k1t = _k1_from_dm(χD, χM, U)
    sp, ch = fdga.susceptibilities(_StandIn(k1a, k1t, U, NP))
    assert sp.shape == shape and ch.shape == shape
    np.testing.assert_allclose(sp, χM, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(ch, χD, rtol=1e-14, atol=1e-14)

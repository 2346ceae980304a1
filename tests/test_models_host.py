"""Model inputs of the host package (src/models/siam.jl:9-29, src/models/hubbard.jl:8-44,
src/dyson.jl:33-41) on the CPU: the Hubbard and SIAM bare Green functions against the oracle's
(which are pinned on special k and the reference's occupation goldens), and compute_occupation
against those goldens (test/test_hubbard.jl:82-88) directly."""
import numpy as np
import pytest

import fdga
import pyoracle as po


@pytest.mark.parametrize("t2,t3", [(0.0, 0.0), (-0.2, 0.05)])
def test_hubbard_bare_green_matches_oracle(t2, t3):
    T, t1, mu, nG, LG = 0.3, 1.0, 0.4, 7, 6
    o = po.Oracle(nonlocal_=True, nG=nG, nK1=2, nK2=(1, 1), nK3=(1, 1), mPinu=1, T=T, U=1.0, LG=LG, L=1)
    o.hubbard(mu, t1, t2, t3)
    ref = o.get(po.GBARE).reshape(2 * nG, LG * LG, order="F")
    got = fdga.hubbard_bare_Green(nG, T, LG, μ=mu, t1=t1, t2=t2, t3=t3)
    assert got.shape == ref.shape
    np.testing.assert_allclose(got, ref, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("D", [None, 3.0])
def test_siam_bare_green_matches_oracle(D):
    T, e, Delta, nG = 0.1, 0.25, 0.7, 9
    o = po.Oracle(nG=nG, nK1=2, nK2=(1, 1), nK3=(1, 1), mPinu=1, T=T, U=1.0)
    o.siam(e, Delta, D)
    ref = o.get(po.GBARE)
    got = fdga.siam_bare_Green(nG, T, e=e, Δ=Delta, D=D)
    np.testing.assert_allclose(got, ref, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("mu,occ", [(-4.0, 0.0502663698543071), (-2.0, 0.2057188296739284), (0.0, 0.5),
                                    (4.0, 1 - 0.0502663698543071), (2.0, 1 - 0.2057188296739284)])
def test_compute_occupation_golden(mu, occ):
    G = fdga.hubbard_bare_Green(20, 0.5, 8, μ=mu, t1=1.0)
    assert abs(fdga.compute_occupation(G, 0.5, 64) - occ) < 1e-12

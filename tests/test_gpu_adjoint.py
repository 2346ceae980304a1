"""Tangent and adjoint of build_K3_cache! followed by BSE_K3! of the three channels
(fdga_bse_k3_jvp / fdga_bse_k3_vjp) on the GPU, as a map of (F, FL.K3, Πpp/Πph) to Fbuff.K3.

Pins:
* the JVP against central differences of the device composite with h = 1e-4 (FP64).  With dΠ = 0
  the composite is quadratic in (F, FL) (the caches are affine in F, the contraction bilinear), so
  the central difference is exact and the agreement is at rounding level; with dΠ the composite is
  cubic and the difference carries its h² term (tolerance 1e-7);
* the JVP against central differences of the oracle's composite (the CPU restatement);
* the VJP against the JVP through <ō, J v> = <J^T ō, v> for random v and ō (pairing sum(u v),
  no conjugation), to rounding.
Cases: the local solver with a local reference vertex and RefVertex tables, and 2 x 2 and 3 x 3
lattices (the latter with unequal K3 boxes).
"""
import numpy as np
import pytest

import pyoracle as po
from helpers import BOX, make_pair, relerr

pytestmark = pytest.mark.gpu
CHS = (po.PCH, po.TCH, po.ACH)
CASES = [("local_vertexref", None), ("nl_L2", None), ("nl_L3", dict(BOX, nK3=(3, 2)))]


def _rand(rng, n, s=0.3):
    return s * (rng.standard_normal(n) + 1j * rng.standard_normal(n))


def _point(S, rng):
    """Random F (K1, K2, K3 of the three channels), FL.K3 and Πpp, Πph of S's sizes."""
    from fdga import _lib as L
    F = {(ch, cls): _rand(rng, S.vertex(0, ch, cls).size) for ch in CHS for cls in (1, 2, 3)}
    FL = {ch: _rand(rng, S.vertex(2, ch, 3).size) for ch in CHS}
    Pi = [_rand(rng, S.size(L.PIPP), 0.1), _rand(rng, S.size(L.PIPH), 0.1)]
    return F, FL, Pi


def _axpy(x, v, h):
    F, FL, Pi = x
    dF, dFL, dPi = v
    return ({k: F[k] + h * dF[k] for k in F}, {k: FL[k] + h * dFL[k] for k in FL}, [a + h * b for a, b in zip(Pi, dPi)])


def _load(S, x):
    from fdga import _lib as L
    F, FL, Pi = x
    for (ch, cls), a in F.items():
        S.set_vertex(0, ch, cls, a)
    for ch, a in FL.items():
        S.set_vertex(2, ch, 3, a)
    S.set(L.PIPP, Pi[0])
    S.set(L.PIPH, Pi[1])


def _composite(S, x):
    """build_K3_cache!, then BSE_K3! for p, a, t at x; Fbuff.K3 of the three channels."""
    import fdga
    _load(S, x)
    fdga.build_K3_cache(S)
    for ch in (po.PCH, po.ACH, po.TCH):
        fdga.BSE_K3(S, ch)
    return np.concatenate([S.vertex(1, ch, 3) for ch in CHS])


def _composite_oracle(O, x):
    F, FL, Pi = x
    for (ch, cls), a in F.items():
        O.set_vertex(po.VTX_F, ch, cls, a)
    for ch, a in FL.items():
        O.set_vertex(po.VTX_FL, ch, 3, a)
    O.set(po.PIPP, Pi[0])
    O.set(po.PIPH, Pi[1])
    O.build_k3_cache()
    for ch in (po.PCH, po.ACH, po.TCH):
        O.bse_k3(ch)
    return np.concatenate([O.vertex(po.VTX_FBUFF, ch, 3) for ch in CHS])


def _jvp(S, Tn, v):
    import fdga
    _load(Tn, v)
    fdga.BSE_K3_jvp(S, Tn)
    return np.concatenate([Tn.vertex(1, ch, 3) for ch in CHS])


@pytest.mark.parametrize("case,box", CASES, ids=[c for c, _ in CASES])
def test_jvp_matches_central_differences(case, box):
    import fdga
    S, O = make_pair(case, seed=21, box=box)
    Tn = fdga.tangent_solver(S)
    rng = np.random.default_rng(5)
    x = _point(S, rng)
    v = _point(S, rng)
    h = 1e-4
    # quadratic direction (dΠ = 0): exact central difference
    v0 = (v[0], v[1], [np.zeros_like(a) for a in v[2]])
    fd0 = (_composite(S, _axpy(x, v0, h)) - _composite(S, _axpy(x, v0, -h))) / (2 * h)
    _load(S, x)
    j0 = _jvp(S, Tn, v0)
    assert np.abs(j0).max() > 1e-3
    assert relerr(j0, fd0) < 1e-9, f"{case} dΠ = 0: {relerr(j0, fd0):.3e}"
    # full direction
    fd = (_composite(S, _axpy(x, v, h)) - _composite(S, _axpy(x, v, -h))) / (2 * h)
    _load(S, x)
    j = _jvp(S, Tn, v)
    assert relerr(j, fd) < 1e-7, f"{case}: {relerr(j, fd):.3e}"
    # the oracle's composite, differenced the same way
    fdo = (_composite_oracle(O, _axpy(x, v, h)) - _composite_oracle(O, _axpy(x, v, -h))) / (2 * h)
    assert relerr(j, fdo) < 1e-7, f"{case} against the oracle: {relerr(j, fdo):.3e}"
    Tn.close()
    S.close()


@pytest.mark.parametrize("case,box", CASES, ids=[c for c, _ in CASES])
def test_vjp_is_the_transpose_of_the_jvp(case, box):
    import fdga
    from fdga import _lib as L
    S, _ = make_pair(case, seed=22, box=box)
    Tn = fdga.tangent_solver(S)
    rng = np.random.default_rng(6)
    x = _point(S, rng)
    v = _point(S, rng)
    _load(S, x)
    jv = _jvp(S, Tn, v)
    ob = _rand(rng, jv.size)
    n3 = jv.size // 3
    for k, ch in enumerate(CHS):
        Tn.set_vertex(1, ch, 3, ob[k * n3:(k + 1) * n3])
    fdga.BSE_K3_vjp(S, Tn)
    lhs = np.sum(ob * jv)
    dF, dFL, dPi = v
    rhs = sum(np.sum(Tn.vertex(0, ch, cls) * dF[ch, cls]) for ch in CHS for cls in (1, 2, 3))
    rhs += sum(np.sum(Tn.vertex(2, ch, 3) * dFL[ch]) for ch in CHS)
    rhs += np.sum(Tn.get(L.PIPP) * dPi[0]) + np.sum(Tn.get(L.PIPH) * dPi[1])
    scale = np.sum(np.abs(ob * jv))
    assert abs(lhs - rhs) < 1e-12 * scale, f"{case}: <ō, J v> = {lhs}, <J^T ō, v> = {rhs}"
    # every input receives a cotangent
    for ch in CHS:
        for cls in (1, 2, 3):
            assert np.abs(Tn.vertex(0, ch, cls)).max() > 0, (ch, cls)
        assert np.abs(Tn.vertex(2, ch, 3)).max() > 0
    assert np.abs(Tn.get(L.PIPP)).max() > 0 and np.abs(Tn.get(L.PIPH)).max() > 0
    Tn.close()
    S.close()


def test_tangent_context_errors():
    import fdga
    S, _ = make_pair("nl_L2", seed=1)
    bad = fdga.Solver(nG=S.nG, nK1=S.nK1, nK2=S.nK2, nK3=S.nK3, T=S.T, U=1.0, mPinu=S.mPinu, nonlocal_=True,
                      LG=S.LG, L_=S.L)
    with pytest.raises(fdga.FdgaError, match="U = 0"):
        fdga.BSE_K3_jvp(S, bad)
    other = fdga.Solver(nG=S.nG, nK1=S.nK1, nK2=S.nK2, nK3=(2, 2), T=S.T, U=0.0, mPinu=S.mPinu, nonlocal_=True,
                        LG=S.LG, L_=S.L)
    with pytest.raises(fdga.FdgaError, match="dimensions"):
        fdga.BSE_K3_vjp(S, other)
    S32, _ = make_pair("nl_L2", seed=1, precision="fp32")
    with pytest.raises(fdga.FdgaError, match="FP64"):
        fdga.BSE_K3_jvp(S32, fdga.tangent_solver(S32))
    L, _ = make_pair("local_ref", seed=1)
    L.init_sym_grp()
    with pytest.raises(fdga.FdgaError, match="symmetry"):
        fdga.BSE_K3_vjp(L, fdga.tangent_solver(L))
    for X in (bad, other, S, S32, L):
        X.close()

"""Pins of the CPU oracle against the reference's own golden vectors and invariants.

Reference fixtures used here:
  * test/test_siam_scPA.jl:28-31   Σ and K1 (a, p, t) of the half-filled SIAM scPA solve
  * test/test_siam_scPA.jl:60-63   Σ of the doped SIAM scPA solve
  * test/test_siam_fdPA.jl:29-33   fdPA from the trivial reference ≡ scPA to 1e-10
  * test/test_siam_fdPA.jl:52-68   fdPA from a shifted reference tracks scPA (tolerances)
  * test/test_siam_fdPA.jl:84      golden Σ(πT) of the box-8 fdPA solve

Finding (recorded in the design notes): the goldens are reproduced only with the two
conventions they were produced with, which differ from the current reference source:
  1. the bubble ν-mesh equals the K1 mesh (mΠν_factor = 1; src/ParquetSolver.jl:92 now
     defaults to 6), and
  2. the fdPA SDE without the extra "- (n0 - 1/2) U i" line (src/SDE.jl:26-31).
With those two conventions every golden is met to <= 3e-4 relative (the remaining gap is the
NLsolve Anderson stopping point at ftol = 1e-4 and MatsubaraFunctions' class representatives,
both un-vendored), and test_current_defaults_miss_goldens shows the size of the deviation
with the current defaults (7 % on K1, 26 % on the fdPA Σ), far above that gap.
"""
import numpy as np
import pytest

import pyoracle as po

T = 0.1
U = 1.0


def interp_bosonic(K, nK1, xs):
    ms = np.arange(-(nK1 - 1), nK1)
    vals = 2 * np.pi * T * ms
    return np.interp(xs, vals, K.real) + 1j * np.interp(xs, vals, K.imag)


# `S.F.γa.K1.(2π * T .* -2:2)` parses as the range (-4πT):1:2 (4 points, 3 of them off-mesh);
# MeshFunction evaluation at a Float64 is taken as linear interpolation on the mesh.
XS = np.arange(2 * np.pi * T * -2, 2 + 1e-12, 1.0)


def scpa_solver(e, Delta, D, nmax, nK2, nK3, mp=1):
    nG, nK1 = 6 * nmax, 4 * nmax
    o = po.Oracle(nG=nG, nK1=nK1, nK2=nK2, nK3=nK3, mPinu=mp, T=T, U=U)
    o.siam(e, Delta, D)
    o.init()
    o.init_sym_grp()
    return o, nG, nK1


def test_siam_scpa_halffilling_golden():
    o, nG, nK1 = scpa_solver(0.0, np.pi / 5, 10.0, 6, (6, 6), (6, 6))
    it, res = o.solve(po.SCPA, maxiter=100, tol=1e-4, beta=0.85, mem=8)
    assert res < 1e-4
    S = o.get(po.SIGMA)[[n + nG for n in (-2, -1, 0, 1)]]
    gS = [-0.052138235296134906, -0.03838544776344314, 0.03838544776344314, 0.052138235296134906]
    np.testing.assert_allclose(S, gS, rtol=3e-4)
    gold = {
        po.ACH: [0.13203850929270397, 0.5403615530152339, 0.2333246221064017, 0.09056300899983459],
        po.PCH: [-0.10420799999591804, -0.2403951910434166, -0.15592452265704748, -0.07622568434721624],
        po.TCH: [0.013898648018808482, 0.1499562726081748, 0.03867632419082161, 0.007160400841240708],
    }
    for ch, g in gold.items():
        K1 = o.vertex(po.VTX_F, ch, 1)
        np.testing.assert_allclose(interp_bosonic(K1, nK1, XS), g, rtol=3e-4)
    # particle-hole symmetry at half filling: iΣ real and odd
    Sall = o.get(po.SIGMA)
    assert np.max(np.abs(Sall.imag)) < 1e-12
    np.testing.assert_allclose(Sall, -Sall[::-1], atol=1e-12)


def test_siam_scpa_doped_golden():
    o, nG, nK1 = scpa_solver(0.5, np.pi / 5, 10.0, 6, (7, 6), (7, 6))
    o.solve(po.SCPA, maxiter=100, tol=1e-4, beta=0.85, mem=8)
    S = o.get(po.SIGMA)[[n + nG for n in (-2, -1, 0, 1)]]
    gS = [-0.0389123277075552 - 0.16855090184215607j, -0.025252640312580586 - 0.17429637478745583j,
          0.025252640312580586 - 0.17429637478745583j, 0.0389123277075552 - 0.16855090184215607j]
    np.testing.assert_allclose(S, gS, rtol=1e-4)
    gold = {
        po.ACH: [0.11925962005661812 + 8.57514999021054e-5j, 0.416232811242488 + 3.319936929625957e-5j,
                 0.20353141073439696 - 8.209974062547027e-5j, 0.08259294412067451 - 7.660306021755952e-5j],
        po.PCH: [-0.12570450372739117 + 0.06583917638195431j, -0.24548654160724023 + 0.021014183409764874j,
                 -0.17578586892780296 - 0.05408344507941078j, -0.09544981624337806 - 0.06686768343644132j],
        po.TCH: [0.011016969129875598 + 7.455601769977568e-5j, 0.09533547032272821 + 2.4477973720973318e-5j,
                 0.028843799251846686 - 6.260706942592786e-5j, 0.005828299701446311 - 7.32415771550901e-5j],
    }
    for ch, g in gold.items():
        K1 = interp_bosonic(o.vertex(po.VTX_F, ch, 1), nK1, XS)
        # absolute scale: the tiny imaginary parts are compared against the magnitude of K1
        np.testing.assert_allclose(K1, g, atol=3e-4 * np.max(np.abs(g)))


def _fdpa_setup(mp, fdpa_h0, nmax=6):
    D, e, Delta = 10.0, -0.3, np.pi / 3
    nG, nK1 = 8 * nmax, 8 * nmax

    def mk():
        o = po.Oracle(nG=nG, nK1=nK1, nK2=(nmax, nmax), nK3=(nmax, nmax), mPinu=mp, T=T, U=U, fdpa_h0=fdpa_h0)
        o.siam(e, Delta, D)
        o.init()
        o.init_sym_grp()
        return o
    return mk


def test_siam_fdpa_trivial_reference_is_scpa():
    """test/test_siam_fdPA.jl:29-33 (fdPA from G0 = Σ0 = 0, F0 = U equals scPA, < 1e-10)."""
    mk = _fdpa_setup(1, False)
    a, b = mk(), mk()
    a.solve(po.SCPA)
    b.solve(po.FDPA)
    assert np.max(np.abs(a.get(po.SIGMA) - b.get(po.SIGMA))) < 1e-10
    assert np.max(np.abs(a.flatten() - b.flatten())) < 1e-10


def _shifted_reference(mp, fdpa_h0, nmax_ref=12, nmax_fd=12):
    D, e, Delta = 10.0, -0.3, np.pi / 3
    nG, nK1 = 8 * nmax_ref, 8 * nmax_ref
    s0 = po.Oracle(nG=nG, nK1=nK1, nK2=(nmax_ref,) * 2, nK3=(nmax_ref,) * 2, mPinu=mp, T=T, U=U)
    s0.siam(e, Delta, D)
    s0.init()
    s0.init_sym_grp()
    s0.solve(po.SCPA)
    nK1f = 8 * nmax_fd
    s = po.Oracle(nG=nG, nK1=nK1f, nK2=(nmax_fd,) * 2, nK3=(nmax_fd,) * 2, mPinu=mp, T=T, U=U,
                  ref_kind=1, ref_boxes=(nK1, (nmax_ref,) * 2, (nmax_ref,) * 2), fdpa_h0=fdpa_h0)
    s.siam(0.5, np.pi / 5, 20.0)
    s.set(po.G0, s0.get(po.G))
    s.set(po.SIGMA0, s0.get(po.SIGMA))
    for ch in range(3):
        for cls in (1, 2, 3):
            s.set_vertex(po.VTX_F0, ch, cls, s0.vertex(po.VTX_F, ch, cls))
    s.init()
    s.init_sym_grp()
    s.solve(po.FDPA)
    return s0, s, nG


@pytest.mark.slow
def test_siam_fdpa_shifted_reference_tracks_scpa():
    """test/test_siam_fdPA.jl:36-68: fdPA from the scPA solution at (e=-0.3, Δ=π/3, D=10)
    to the target (e=0.5, Δ=π/5, D=20) agrees with the target scPA solve."""
    s0, s, nG = _shifted_reference(1, False)
    st = po.Oracle(nG=nG, nK1=96, nK2=(12, 12), nK3=(12, 12), mPinu=1, T=T, U=U)
    st.siam(0.5, np.pi / 5, 20.0)
    st.init()
    st.init_sym_grp()
    st.solve(po.SCPA)
    assert np.max(np.abs(s.get(po.SIGMA) - st.get(po.SIGMA))) < 3e-5
    tol = {1: 5e-4, 2: 1e-3, 3: 1e-3}
    for ch in range(3):
        for cls in (1, 2, 3):
            d = s.vertex(po.VTX_F, ch, cls) + s0.vertex(po.VTX_F, ch, cls) - st.vertex(po.VTX_F, ch, cls)
            assert np.max(np.abs(d)) < tol[cls], (ch, cls)


@pytest.mark.slow
def test_siam_fdpa_box8_golden():
    """test/test_siam_fdPA.jl:74-84: fdPA with box 8 from the box-12 reference, golden Σ(πT)."""
    s0, s, nG = _shifted_reference(1, False, nmax_ref=12, nmax_fd=8)
    v = s.get(po.SIGMA)[nG]
    np.testing.assert_allclose(v, 0.024643001835742997 - 0.17494219707558506j, rtol=2e-4)


def test_current_defaults_miss_goldens():
    """With the current source defaults (mΠν_factor = 6) the scPA goldens are missed by
    percents, which is why the golden tests above run with mΠν_factor = 1."""
    o, nG, nK1 = scpa_solver(0.0, np.pi / 5, 10.0, 6, (6, 6), (6, 6), mp=6)
    o.solve(po.SCPA)
    K1 = interp_bosonic(o.vertex(po.VTX_F, po.ACH, 1), nK1, XS)
    rel = np.abs(K1 - np.array([0.13203850929270397, 0.5403615530152339, 0.2333246221064017,
                                0.09056300899983459])) / 0.54
    assert rel.max() > 1e-2

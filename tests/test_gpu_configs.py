"""The benchmark configurations on the GPU against the oracle or an exact reduction.

  * the small SIAM scPA solve at its exact boxes (nK1 = 33, nK2 = (16, 16), nK3 = (8, 8), D = ∞,
    e = 0, Δ = π/5, U = 1, T = 0.1, mΠν = 6, nG = 96, init_sym_grp!): the GPU solve and the oracle's
    solve (Anderson, m = 8, β = 0.85) to max|R| < 1e-8 take the same number of iterations and agree
    to 1e-9; so do χ_sp and χ_ch;
  * the lattice workload's frequency boxes (nK1 = 64, nK2 = (32, 32), nK3 = (16, 32), mΠν = 32,
    nG = 256, T = 0.2, U = 2) with its real inputs (the half-filled Hubbard dispersion t = 1, the local
    reference system: Gaussian reference K3, momentum-averaged G0) on a 4 x 4 lattice the oracle can
    hold: one fdPA iteration from the reference state and one from a random vertex, FP64 to 1e-10;
  * the 8-point batched sweep at its full boxes (tests/test_gpu_sweep.py).
"""
import os
import sys

import numpy as np
import pytest

import pyoracle as po
from helpers import relerr

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import fdga  # noqa: E402


def test_c0_exact_boxes_solved_to_1e_8():
    import bench
    w = bench.C0
    print("c0: GPU solve", flush=True)
    S = fdga.parquet_solver_siam_parquet_approximation(w["nG"], w["nK1"], w["nK2"], w["nK3"], e=0.0, Δ=w["Delta"], D=None,
                                                       T=w["T"], U=w["U"], mΠν_factor=w["mPinu"])
    S.init_sym_grp()
    it, res = fdga.solve(S, strategy="scPA", tol=1e-8, maxiter=100)
    print(f"c0: GPU {it} iterations, max|f| {res:.2e}; oracle solve", flush=True)
    O = po.Oracle(nG=w["nG"], nK1=w["nK1"], nK2=w["nK2"], nK3=w["nK3"], mPinu=w["mPinu"], T=w["T"], U=w["U"])
    O.siam(0.0, w["Delta"], None)
    O.init()
    O.init_sym_grp()
    it_o, res_o = O.solve(po.SCPA, maxiter=100, tol=1e-8)
    assert res < 1e-8 and res_o < 1e-8
    assert it == it_o, (it, it_o)
    assert relerr(fdga.flatten(S), O.flatten()) < 1e-9
    sp, ch = fdga.susceptibilities(S)
    sp_o, ch_o = fdga.chi_from_k1(O.vertex(po.VTX_F, po.ACH, 1), O.vertex(po.VTX_F, po.TCH, 1), w["U"])
    assert relerr(sp, sp_o) < 1e-9 and relerr(ch, ch_o) < 1e-9
    S.close()


def _c2_pair(L, precision="fp64"):
    import bench
    w = dict(bench.C2, L=L)
    S = bench.build_c2(precision, w)
    O = po.Oracle(nonlocal_=True, nG=w["nG"], nK1=w["nK1"], nK2=w["nK2"], nK3=w["nK3"], mPinu=w["mPinu"], T=w["T"],
                  U=w["U"], LG=L, L=L, ref_kind=1, ref_boxes=(w["nK1"], w["nK2"], w["nK3"]))
    O.hubbard(w["mu"], w["t1"])
    Gb = O.get(po.GBARE).reshape(2 * w["nG"], L * L, order="F")
    O.set(po.G0, np.repeat(Gb.mean(axis=1, keepdims=True), L * L, axis=1).reshape(-1, order="F"))
    for (c, cls), a in bench.c2_reference_vertex(w, seed=0).items():
        O.set_vertex(po.VTX_F0, c, cls, a)
    O.init()
    return S, O


@pytest.mark.timeout(400)
def test_c2_boxes_real_inputs_on_a_4x4_lattice():
    print("c2 L=4: building", flush=True)
    S, O = _c2_pair(4)
    assert relerr(S.get(fdga._lib.PIPP), O.get(po.PIPP)) < 1e-12
    fdga.iterate_solver(S, "fdPA")
    O.iterate(po.FDPA)
    x = O.flatten()
    assert np.max(np.abs(x)) > 1e-2                      # a nontrivial correction vertex
    assert relerr(fdga.flatten(S), x) < 1e-10
    # from a random vertex (every BSE term active)
    rng = np.random.default_rng(5)
    y = x + 0.02 * (rng.standard_normal(x.size) + 1j * rng.standard_normal(x.size))
    fdga.unflatten(S, y)
    O.unflatten(y)
    fdga.iterate_solver(S, "fdPA")
    O.iterate(po.FDPA)
    e = relerr(fdga.flatten(S), O.flatten())
    print(f"c2 L=4: second iteration rel err {e:.2e}", flush=True)
    assert e < 1e-10
    S.close()
